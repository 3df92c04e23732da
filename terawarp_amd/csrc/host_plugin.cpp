// host_plugin.cpp -- the reference's plugin surface over the C ABI.
//
// Mirrors plugin_Bigimagewarp.cpp: menulist/funclist (26-38), dofunc
// (73-107: argument-count check, input/output vectors), printHelp (109-120)
// and bigimagewarp (123-336: marker check and read, subject image read,
// output naming <out>/<basename>_affine|_stps.v3draw, dispatch to the
// affine or TPS driver).  The compute runs on the GPU through terawarp.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/terawarp.h"
#include "../../include/terawarp_host.h"

namespace tw_host {
void set_error(const std::string& m);
}

namespace {

bool fail(const std::string& m) {
    tw_host::set_error(m);
    std::fprintf(stderr, "ERROR: %s\n", m.c_str());
    return false;
}

bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

bool parse_num(const char* s, double* v) {
    if (!s) return false;
    char* end = nullptr;
    *v = std::strtod(s, &end);
    return end != s && *end == '\0' && std::isfinite(*v);
}

const char* const kMenu[] = {"bigimagewarp", "about"};
const char* const kFunc[] = {"bigimagewarp", "help"};

void print_help() {
    std::printf(
        "\nUsage: terawarp -i <input_image_sub> <input_marker_sub> <input_marker_tar> -o <output_folder> -p "
        "warp_mode device_mode resize_mode deformation_mode para_x para_y para_z newsize_x newsize_y newsize_z\n"
        "\t warp_mode:        0 non-rigid, 1 rigid\n"
        "\t device_mode:      0 reference-exact (FP64, GPU), 1 GPU plain form, 2 GPU mixed precision\n"
        "\t resize_mode:      1 to write an output of extent newsize_x/y/z\n"
        "\t deformation_mode: non-rigid 0 affine, 1 TPS, 2 scaling; rigid 0 rotation, 1 translation\n"
        "\t para_x/y/z:       scaling factors (0.1-10), rotation angles (deg, -180..180) or translation steps\n"
        "\t legacy form:      -p program_mode(1 affine, 0 TPS) device_mode downsample_mode x x x x\n");
}

struct Ctx {
    tw_ctx* c = nullptr;
    ~Ctx() {
        if (c) tw_destroy(c);
    }
};

}  // namespace

extern "C" {

int tw_plugin_menu_count(void) { return 2; }
const char* tw_plugin_menu_item(int i) { return (i >= 0 && i < 2) ? kMenu[i] : nullptr; }
int tw_plugin_func_count(void) { return 2; }
const char* tw_plugin_func_item(int i) { return (i >= 0 && i < 2) ? kFunc[i] : nullptr; }
float tw_plugin_version(void) { return 1.1f; }

int tw_plugin_parse_params(const char* const* paras, int n, tw_warp_params* out) {
    if (!out || (!paras && n > 0)) return TW_E_ARG;
    std::memset(out, 0, sizeof *out);
    std::vector<double> v((size_t)std::max(n, 0));
    for (int i = 0; i < n; i++)
        if (!parse_num(paras[i], &v[i])) {
            tw_host::set_error(std::string("parameter ") + std::to_string(i) + " is not a number: " +
                               (paras[i] ? paras[i] : "(null)"));
            return TW_E_ARG;
        }
    if (n == 7) {  // reference code form: program_mode device_mode downsample_mode x x x x
        out->legacy = 1;
        out->kind = v[0] != 0 ? 0 : 1;
        out->device_mode = (int)v[1];
        out->downsample = (int)v[2];
    } else if (n == 10) {
        const int warp_mode = (int)v[0], def = (int)v[3];
        out->device_mode = (int)v[1];
        out->resize = (int)v[2];
        for (int a = 0; a < 3; a++) out->para[a] = v[4 + a];
        for (int a = 0; a < 3; a++) out->newsize[a] = (int64_t)v[7 + a];
        if (warp_mode == 0 && def >= 0 && def <= 2) out->kind = def;  // affine, TPS, scaling
        else if (warp_mode == 1 && (def == 0 || def == 1)) out->kind = 3 + def;  // rotation, translation
        else {
            tw_host::set_error("warp_mode/deformation_mode combination is not defined");
            return TW_E_ARG;
        }
        if (out->resize != 0 && out->resize != 1) {
            tw_host::set_error("resize_mode must be 0 or 1");
            return TW_E_ARG;
        }
        if (out->resize)
            for (int a = 0; a < 3; a++)
                if (out->newsize[a] <= 0) {
                    tw_host::set_error("newsize must be positive when resize_mode = 1");
                    return TW_E_ARG;
                }
        if (out->kind == 2)
            for (int a = 0; a < 3; a++)
                if (!(out->para[a] >= 0.1 && out->para[a] <= 10.0)) {
                    tw_host::set_error("scaling factors must lie in [0.1, 10]");
                    return TW_E_ARG;
                }
        if (out->kind == 3)
            for (int a = 0; a < 3; a++)
                if (!(out->para[a] >= -180.0 && out->para[a] <= 180.0)) {
                    tw_host::set_error("rotation angles must lie in [-180, 180]");
                    return TW_E_ARG;
                }
    } else {
        tw_host::set_error("ERROR: no enough para! (expected 10 values, or the 7-value legacy form)");
        return TW_E_ARG;
    }
    if (out->device_mode < 0 || out->device_mode > 2) {
        tw_host::set_error("device_mode must be 0, 1 or 2");
        return TW_E_ARG;
    }
    return TW_OK;
}

int tw_bigimagewarp(const char* img_path, const char* marker_sub, const char* marker_tar, const char* out_dir,
                    const tw_warp_params* p, char* out_path, int out_cap) {
    if (!img_path || !marker_sub || !marker_tar || !out_dir || !p) return fail("null argument"), TW_E_ARG;
    const std::string img(img_path), msub(marker_sub), mtar(marker_tar);
    const bool needs_markers = p->kind == 0 || p->kind == 1;
    // 1. markers (plugin_Bigimagewarp.cpp:143-154)
    std::vector<float> sub, tar;
    int nsub = 0, ntar = 0;
    if (needs_markers) {
        if (!ends_with(mtar, ".marker") || !ends_with(msub, ".marker"))
            return fail("at least one marker file is invalid (expected .marker files)"), TW_E_ARG;
        nsub = tw_markers_read(marker_sub, nullptr, 0);
        ntar = tw_markers_read(marker_tar, nullptr, 0);
        if (nsub < 0 || ntar < 0) return fail(tw_io_last_error()), TW_E_IO;
        if (nsub == 0 || ntar == 0 || nsub != ntar)
            return fail("target or subject control points is invalid!"), TW_E_ARG;
        sub.resize((size_t)nsub * 3);
        tar.resize((size_t)ntar * 3);
        tw_markers_read(marker_sub, sub.data(), nsub);
        tw_markers_read(marker_tar, tar.data(), ntar);
    }
    // 2. subject image (only the in-memory .v3draw path; 155-210)
    if (!ends_with(img, ".v3draw"))
        return fail("slice-directory (tera-scale) input is not part of this build; pass a .v3draw stack"),
               TW_E_UNSUPPORTED;
    int64_t sz[4];
    int dt = 0, sw = 0, hl = 0;
    if (tw_v3draw_header(img_path, sz, &dt, &sw, &hl)) return fail(tw_io_last_error()), TW_E_IO;
    if (dt != 1 && dt != 2) return fail("only uint8 and uint16 stacks are supported"), TW_E_UNSUPPORTED;
    const int64_t nvox = sz[0] * sz[1] * sz[2] * sz[3];
    std::vector<uint8_t> src((size_t)(nvox * dt));
    if (tw_v3draw_read(img_path, src.data(), (int64_t)src.size())) return fail(tw_io_last_error()), TW_E_IO;
    int64_t osz[4] = {sz[0], sz[1], sz[2], sz[3]};
    if (p->resize)
        for (int a = 0; a < 3; a++) osz[a] = p->newsize[a];
    std::vector<uint8_t> dst((size_t)(osz[0] * osz[1] * osz[2] * osz[3] * dt), 0);
    // Output name: <out>/<basename><suffix>.v3draw (166-175)
    const size_t slash = img.find_last_of('/');
    const size_t b0 = slash == std::string::npos ? 0 : slash + 1;
    const std::string base = img.substr(b0, img.find_last_of('.') - b0);
    static const char* const suffix[] = {"_affine", "_stps", "_scaling", "_rotation", "_translation"};
    const std::string out = std::string(out_dir) + "/" + base + suffix[p->kind] + ".v3draw";

    Ctx ctx;
    int rc = tw_create(0, &ctx.c);
    if (rc) return fail("no usable GPU (terawarp runs on the MI355X only)"), rc;
    tw_volume vs = {}, vd = {};
    vs.data = src.data();
    vs.nx = sz[0]; vs.ny = sz[1]; vs.nz = sz[2]; vs.nc = sz[3];
    vs.dtype = dt == 1 ? TW_U8 : TW_U16;
    vs.mem = TW_MEM_HOST;
    vd = vs;
    vd.data = dst.data();
    vd.nx = osz[0]; vd.ny = osz[1]; vd.nz = osz[2];
    const int precision = p->device_mode == 0 ? TW_EXACT : (p->device_mode == 1 ? TW_PLAIN : TW_MIXED);
    if (p->kind == 1) {
        // imgwarp_smallmemory_image(sub, tar, 4, 4, 4, ...): kernel centres are
        // the target-space markers mapped onto the subject markers (197).
        rc = tw_warp_tps(ctx.c, &vs, &vd, tar.data(), sub.data(), nsub, 4, 0.2, TW_INTERP_LINEAR, precision);
    } else {
        double M[16];
        if (p->kind == 0) rc = tw_affine_fit(tar.data(), sub.data(), nsub, M);  // target -> subject (1351)
        else rc = tw_mode_matrix(p->kind == 2 ? 2 : (p->kind == 3 ? 0 : 1), p->para, sz[0], sz[1], sz[2], M);
        if (rc) return fail("cannot build the transform matrix (degenerate markers or parameters)"), rc;
        rc = tw_resample_affine(ctx.c, &vs, &vd, M, TW_INTERP_LINEAR);
    }
    if (rc) return fail(std::string("warp failed: ") + tw_last_error(ctx.c)), rc;
    if (tw_v3draw_write(out.c_str(), dst.data(), osz, dt)) return fail(tw_io_last_error()), TW_E_IO;
    if (out_path && out_cap > 0) {
        std::strncpy(out_path, out.c_str(), (size_t)out_cap - 1);
        out_path[out_cap - 1] = '\0';
    }
    return TW_OK;
}

int tw_plugin_dofunc(const char* func_name, const char* const* inputs, int n_inputs, const char* const* paras,
                     int n_paras, const char* const* outputs, int n_outputs) {
    const std::string fn = func_name ? func_name : "";
    if (fn != "bigimagewarp") {
        print_help();
        return fn == "help" ? 1 : 0;
    }
    // plugin_Bigimagewarp.cpp:78: input = {files, paras}, output = {folder}
    if (n_inputs != 3 || n_outputs != 1 || !inputs || !outputs)
        return fail("ERROR: no enough para! (need image, subject marker, target marker and an output folder)");
    tw_warp_params p;
    if (tw_plugin_parse_params(paras, n_paras, &p) != TW_OK) return fail(tw_io_last_error());
    return tw_bigimagewarp(inputs[0], inputs[1], inputs[2], outputs[0], &p, nullptr, 0) == TW_OK ? 1 : 0;
}

}  // extern "C"
