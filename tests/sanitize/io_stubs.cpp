// The library entry points the file readers call, for the sanitizer build of
// scene_io.cpp / image_io.cpp alone: the error-string pair they report
// through, and scene calls that refuse (the sanitized tests load files
// through hh_gltf_load's sink and hh_ibl_from_hdr, which never reach them).
#include <string>

#include "helios_hip.h"

static thread_local std::string g_msg;

hh_status hh_internal_fail(hh_status s, const char* msg) {
    g_msg = msg ? msg : "";
    return s;
}

extern "C" {
const char* hh_last_error(void) { return g_msg.c_str(); }
hh_status hh_scene_add_texture(hh_scene, uint32_t, uint32_t, uint32_t, const void*, uint32_t*) {
    return HH_ERR_INVALID_ARG;
}
hh_status hh_scene_add_material(hh_scene, const hh_material_desc*, uint32_t*) { return HH_ERR_INVALID_ARG; }
hh_status hh_scene_add_mesh(hh_scene, const hh_vertex*, uint32_t, const uint32_t*, uint32_t, const hh_submesh*,
                            uint32_t, const uint32_t*, uint32_t, uint32_t*) {
    return HH_ERR_INVALID_ARG;
}
hh_status hh_scene_add_instance(hh_scene, const hh_instance_desc*, uint32_t*) { return HH_ERR_INVALID_ARG; }
hh_status hh_scene_add_light(hh_scene, const hh_light_desc*, uint32_t*) { return HH_ERR_INVALID_ARG; }
hh_status hh_scene_set_ibl(hh_scene, uint32_t, const float*) { return HH_ERR_INVALID_ARG; }
}
