"""Regenerate ``hosek_rgb.bin``: the Hosek-Wilkie RGB sky coefficient table.

The table is published data from "An Analytic Model for Full Spectral Sky-Dome
Radiance" (Hosek & Wilkie, SIGGRAPH 2012) and "Adding a Solar Radiance Function
to the Hosek Skylight Model" (IEEE CG&A 2013), RGB dataset version 1.4a.  The
reference ships it as a C array include (include/gfx/hosek_data_rgb.inl:76-3832);
this script reads the numbers out of that file and stores them as raw
little-endian float64 in the order

    datasetRGB1, datasetRGB2, datasetRGB3        (1080 doubles each)
    datasetRGBRad1, datasetRGBRad2, datasetRGBRad3 (120 doubles each)

i.e. 3600 doubles.  Layout inside a datasetRGB block: [albedo 0|1][turbidity
1..10][6 spline control points][9 coefficients A..I]; datasetRGBRad blocks use
one coefficient.  See HOSEK_ATTRIBUTION.txt for the data licence notice.

Usage: python make_hosek_table.py /path/to/hosek_data_rgb.inl
"""
import re
import struct
import sys
from pathlib import Path

ORDER = ["datasetRGB1", "datasetRGB2", "datasetRGB3",
         "datasetRGBRad1", "datasetRGBRad2", "datasetRGBRad3"]
SIZES = {"datasetRGB": 1080, "datasetRGBRad": 120}


def parse(text):
    arrays = {}
    for m in re.finditer(r"double\s+(datasetRGB(?:Rad)?[123])\s*\[\]\s*=\s*\{(.*?)\};", text, re.S):
        body = re.sub(r"//[^\n]*", "", m.group(2))
        vals = [float(tok) for tok in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", body)]
        arrays[m.group(1)] = vals
    return arrays


def main(src):
    arrays = parse(Path(src).read_text())
    out = []
    for name in ORDER:
        vals = arrays[name]
        want = SIZES["datasetRGBRad" if "Rad" in name else "datasetRGB"]
        if len(vals) != want:
            raise SystemExit(f"{name}: expected {want} values, found {len(vals)}")
        out.extend(vals)
    dst = Path(__file__).with_name("hosek_rgb.bin")
    dst.write_bytes(struct.pack("<%dd" % len(out), *out))
    print(f"wrote {dst} ({len(out)} doubles)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/include/gfx/hosek_data_rgb.inl")
