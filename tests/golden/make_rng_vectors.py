"""Writes rng_vectors.json: per-pixel RNG states after rng_init and the first
draws, for pixels and frame indices used by the workloads. Produced by the
oracle's transcription of random.glsl:11-50 (integer-exact); the GPU kernels
are checked against the same values through primary-ray parity."""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
import oracle  # noqa: E402

CASES = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (255, 255, 63), (1919, 1079, 12345), (37, 1000, 4095), (65535, 65535, 0xFFFFFFFF)]

out = []
for x, y, f in CASES:
    r = oracle.RNG(x, y, f)
    st = [int(v) for v in r.s]
    out.append(dict(x=x, y=y, frame=f, state=st, draws=[r.next() for _ in range(8)]))
Path(__file__).with_name("rng_vectors.json").write_text(json.dumps(out, indent=1) + "\n")
print("wrote", len(out), "cases")
