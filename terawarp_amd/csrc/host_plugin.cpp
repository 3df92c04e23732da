// host_plugin.cpp -- the reference's plugin surface over the C ABI.
//
// Mirrors plugin_Bigimagewarp.cpp: menulist/funclist (26-38), dofunc
// (73-107: argument-count check, input/output vectors), printHelp (109-120)
// and bigimagewarp (123-336: marker check and read, subject image read,
// output naming <out>/<basename>_affine|_stps.v3draw, dispatch to the
// affine or TPS driver).  The compute runs on the GPU through terawarp.h.
#include <sys/stat.h>

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

struct SliceDir {
    tw_slice_dir* d = nullptr;
    ~SliceDir() {
        if (d) tw_slice_dir_close(d);
    }
};

// The tera drivers step their output blocks by per_block_siza = 1000 voxels
// and clip the last block of each axis: the TPS warp
// (q_imgwarp_tps_quicksmallmemory.cpp:1264, 1309-1317), the affine warp
// (q_warp_affine_tps.cpp:902, 966-974) and the downsample pass (1531,
// 1595-1602).  (The block-size arrays start at 100 or 500, 1256 / 896 /
// 1521, and are overwritten from per_block_siza before use.)
constexpr int kTeraBlock = 1000;
// Both tera warps multiply the marker coordinates by 4 before fitting
// (x_para, q_imgwarp_tps_quicksmallmemory.cpp:1194-1204; x_power,
// q_warp_affine_tps.cpp:866-872): the markers are placed on a view
// downsampled 4x per axis.  The in-memory warps use them as they are.
constexpr float kTeraMarkerScale = 4.0f;

// Affine / rigid warp of a slice directory block by block
// (q_imagewarp_affine_cyf, q_warp_affine_tps.cpp:851-1329): each output
// block reads the source window that its image under M spans (the bounding
// box of the eight transformed corners, one voxel of margin for the
// rounding of the positions), resamples it on the GPU and is written as a
// block file.
int affine_blocks(tw_ctx* ctx, tw_slice_dir* sd, const double* M, int64_t nx, int64_t ny, int64_t nz, int32_t dtype,
                  const std::string& base) {
    const int bpv = dtype == TW_U8 ? 1 : 2;
    const tw_block_files bf = {base.c_str(), dtype, 0};
    const int64_t nbx = (nx + kTeraBlock - 1) / kTeraBlock, nby = (ny + kTeraBlock - 1) / kTeraBlock,
                  nbz = (nz + kTeraBlock - 1) / kTeraBlock;
    const int64_t full[3] = {nx, ny, nz};
    std::vector<uint8_t> win, out;
    int64_t brick = 0;
    for (int64_t bz = 0; bz < nbz; bz++)
        for (int64_t by = 0; by < nby; by++)
            for (int64_t bx = 0; bx < nbx; bx++, brick++) {
                const int64_t o[3] = {bx * kTeraBlock, by * kTeraBlock, bz * kTeraBlock};
                const int64_t n[3] = {std::min<int64_t>(kTeraBlock, nx - o[0]), std::min<int64_t>(kTeraBlock, ny - o[1]),
                                      std::min<int64_t>(kTeraBlock, nz - o[2])};
                double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                for (int c = 0; c < 8; c++) {
                    const double p[3] = {(double)(o[0] + ((c & 1) ? n[0] - 1 : 0)),
                                         (double)(o[1] + ((c & 2) ? n[1] - 1 : 0)),
                                         (double)(o[2] + ((c & 4) ? n[2] - 1 : 0))};
                    for (int a = 0; a < 3; a++) {
                        const double q = M[4 * a] * p[0] + M[4 * a + 1] * p[1] + M[4 * a + 2] * p[2] + M[4 * a + 3];
                        lo[a] = std::min(lo[a], q);
                        hi[a] = std::max(hi[a], q);
                    }
                }
                int64_t w0[3], wn[3];
                bool empty = false;
                for (int a = 0; a < 3; a++) {
                    const double l = std::floor(lo[a]) - 1, h = std::ceil(hi[a]) + 2;  // [l, h)
                    w0[a] = (int64_t)std::max(0.0, std::min(l, (double)full[a]));
                    const int64_t e = (int64_t)std::max(0.0, std::min(h, (double)full[a]));
                    wn[a] = e - w0[a];
                    empty |= wn[a] < 1;
                }
                out.assign((size_t)(n[0] * n[1] * n[2] * bpv), 0);
                if (!empty) {
                    win.resize((size_t)(wn[0] * wn[1] * wn[2] * bpv));
                    int rc = tw_slice_dir_read(sd, w0[0], w0[1], w0[2], wn[0], wn[1], wn[2], win.data());
                    if (rc) return fail(std::string("slice read failed: ") + tw_tera_last_error()), rc;
                    tw_volume vs = {}, vd = {};
                    vs.data = win.data();
                    vs.nx = wn[0]; vs.ny = wn[1]; vs.nz = wn[2]; vs.nc = 1;
                    vs.x0 = w0[0]; vs.y0 = w0[1]; vs.z0 = w0[2];
                    vs.full_nx = nx; vs.full_ny = ny; vs.full_nz = nz;
                    vs.dtype = dtype;
                    vs.mem = TW_MEM_HOST;
                    vd = vs;
                    vd.data = out.data();
                    vd.nx = n[0]; vd.ny = n[1]; vd.nz = n[2];
                    vd.x0 = o[0]; vd.y0 = o[1]; vd.z0 = o[2];
                    vd.full_nx = vd.full_ny = vd.full_nz = 0;
                    if ((rc = tw_resample_affine(ctx, &vs, &vd, M, TW_INTERP_LINEAR)))
                        return fail(std::string("warp failed: ") + tw_last_error(ctx)), rc;
                }
                int rc = tw_block_file_write((void*)&bf, brick, o[0], o[1], o[2], n[0], n[1], n[2], out.data());
                if (rc) return fail(tw_tera_last_error()), rc;
            }
    return TW_OK;
}

// Preview pass (q_imagewarp_affine_cyf_downsample, q_warp_affine_tps.cpp:
// 1474-2029): the warped slices of <base>_slice downsampled by the factor f
// of tw_downsample_factor to extent n / f, written as block files on the
// 1000-voxel lattice and slices under <base>_downsample.  The reference estimates the matrix by
// its DLT from the target markers and their f-scaled copies, which is
// diag(f, f, f) up to ~1e-12 rounding; the exact matrix is used here, so an
// output voxel is the warped voxel at f times its coordinates.  The slices
// f*k are gathered slab by slab and resampled on the GPU with
// diag(f, f, 1).
int downsample_pass(tw_ctx* ctx, const std::string& base, int64_t nx, int64_t ny, int64_t nz, int32_t dtype) {
    const int f = tw_downsample_factor(nx, ny, nz);
    if (f < 1) return fail("downsample: the largest extent is below 300 voxels (factor 0)"), TW_E_ARG;
    const int64_t onx = nx / f, ony = ny / f, onz = nz / f;
    if (onx < 1 || ony < 1 || onz < 1) return fail("downsample: an extent is below the factor"), TW_E_ARG;
    SliceDir sd;
    int rc = tw_slice_dir_open((base + "_slice").c_str(), &sd.d);
    if (rc) return fail(tw_tera_last_error()), rc;
    int64_t snx, sny, snz;
    int32_t sdt;
    tw_slice_dir_size(sd.d, &snx, &sny, &snz, &sdt);
    if (snx != nx || sny != ny || snz != nz || sdt != dtype) return fail("downsample: warped slices differ"), TW_E_IO;
    const int bpv = dtype == TW_U8 ? 1 : 2;
    std::vector<uint8_t> vol((size_t)(onx * ony * onz * bpv)), slab;
    const int64_t ks_max = std::max<int64_t>(1, std::min<int64_t>(onz, (1ll << 30) / (nx * ny * bpv)));
    const double M[12] = {(double)f, 0, 0, 0, 0, (double)f, 0, 0, 0, 0, 1, 0};
    for (int64_t k0 = 0; k0 < onz; k0 += ks_max) {
        const int64_t ks = std::min(ks_max, onz - k0);
        slab.resize((size_t)(nx * ny * ks * bpv));
        for (int64_t k = 0; k < ks; k++)
            if ((rc = tw_slice_dir_read(sd.d, 0, 0, (k0 + k) * f, nx, ny, 1, slab.data() + (size_t)(k * nx * ny * bpv))))
                return fail(tw_tera_last_error()), rc;
        tw_volume vs = {}, vd = {};
        vs.data = slab.data();
        vs.nx = nx; vs.ny = ny; vs.nz = ks; vs.nc = 1;
        vs.dtype = dtype;
        vs.mem = TW_MEM_HOST;
        vd = vs;
        vd.data = vol.data() + (size_t)(k0 * onx * ony * bpv);
        vd.nx = onx; vd.ny = ony; vd.nz = ks;
        if ((rc = tw_resample_affine(ctx, &vs, &vd, M, TW_INTERP_LINEAR)))
            return fail(std::string("downsample failed: ") + tw_last_error(ctx)), rc;
    }
    const std::string dbase = base + "_downsample";
    const tw_block_files bf = {dbase.c_str(), dtype, 0};
    std::vector<uint8_t> blk;
    int64_t brick = 0;
    for (int64_t z0 = 0; z0 < onz; z0 += kTeraBlock)
        for (int64_t y0 = 0; y0 < ony; y0 += kTeraBlock)
            for (int64_t x0 = 0; x0 < onx; x0 += kTeraBlock, brick++) {
                const int64_t bnx = std::min<int64_t>(kTeraBlock, onx - x0), bny = std::min<int64_t>(kTeraBlock, ony - y0),
                              bnz = std::min<int64_t>(kTeraBlock, onz - z0);
                blk.resize((size_t)(bnx * bny * bnz * bpv));
                for (int64_t z = 0; z < bnz; z++)
                    for (int64_t y = 0; y < bny; y++)
                        std::memcpy(blk.data() + (size_t)(((z * bny) + y) * bnx * bpv),
                                    vol.data() + (size_t)((((z0 + z) * ony + (y0 + y)) * onx + x0) * bpv),
                                    (size_t)(bnx * bpv));
                if ((rc = tw_block_file_write((void*)&bf, brick, x0, y0, z0, bnx, bny, bnz, blk.data())))
                    return fail(tw_tera_last_error()), rc;
            }
    if ((rc = tw_stitch_slices(dbase.c_str(), onx, ony, onz, kTeraBlock, dtype))) return fail(tw_tera_last_error()), rc;
    return TW_OK;
}

// Tera-scale branch of bigimagewarp (plugin_Bigimagewarp.cpp:211-330): the
// subject is a directory of slices; the warp writes <out>/<dir><suffix>_block
// files and <out>/<dir><suffix>_slice slices, then, with downsample_mode,
// the preview pass.  TPS runs through the streaming driver (one fit, one
// grid, exact per-brick source windows, bricks = the reference's blocks,
// positions in its block-local order); affine and rigid warps block by
// block.  The markers of both fitted warps are scaled by kTeraMarkerScale.
int tera_branch(const std::string& dir, const char* out_dir, const tw_warp_params* p, std::vector<float> tar,
                std::vector<float> sub, int nm, const char* suffix, char* out_path, int out_cap) {
    for (float& v : tar) v *= kTeraMarkerScale;
    for (float& v : sub) v *= kTeraMarkerScale;
    SliceDir sd;
    int rc = tw_slice_dir_open(dir.c_str(), &sd.d);
    if (rc) return fail(tw_tera_last_error()), rc;
    int64_t nx, ny, nz;
    int32_t dtype;
    tw_slice_dir_size(sd.d, &nx, &ny, &nz, &dtype);
    std::string name = dir;
    while (name.size() > 1 && name.back() == '/') name.pop_back();
    name = name.substr(name.find_last_of('/') == std::string::npos ? 0 : name.find_last_of('/') + 1);
    const std::string base = std::string(out_dir) + "/" + name + suffix;
    Ctx ctx;
    if ((rc = tw_create(0, &ctx.c))) return fail("no usable GPU (terawarp runs on the MI355X only)"), rc;
    const int precision = p->device_mode == 0 ? TW_EXACT : (p->device_mode == 1 ? TW_PLAIN : TW_MIXED);
    if (p->kind == 1) {
        // imgwarp_smallmemory_CYF(sub, tar, 4, 4, 4, ...) (294): centres are the
        // target-space markers.
        const tw_block_files bf = {base.c_str(), dtype, 0};
        tw_stream_desc d = {};
        d.nx = nx; d.ny = ny; d.nz = nz;
        d.dtype = dtype;
        d.brick = kTeraBlock;
        d.f = 4;
        d.interp = TW_INTERP_LINEAR;
        d.precision = precision;
        d.threads = 8;
        d.flags = TW_STREAM_BLOCK_ORIGIN;
        d.K = nm;
        d.lambda = 0.2;
        d.centres = tar.data();
        d.targets = sub.data();
        d.read = tw_slice_dir_read;
        d.read_user = sd.d;
        d.write = tw_block_file_write;
        d.write_user = (void*)&bf;
        if ((rc = tw_warp_stream(ctx.c, &d, nullptr))) return fail(std::string("warp failed: ") + tw_last_error(ctx.c)), rc;
    } else {
        double M[16];
        if (p->kind == 0) rc = tw_affine_fit(tar.data(), sub.data(), nm, M);
        else rc = tw_mode_matrix(p->kind == 2 ? 2 : (p->kind == 3 ? 0 : 1), p->para, nx, ny, nz, M);
        if (rc) return fail("cannot build the transform matrix (degenerate markers or parameters)"), rc;
        if ((rc = affine_blocks(ctx.c, sd.d, M, nx, ny, nz, dtype, base))) return rc;
    }
    if ((rc = tw_stitch_slices(base.c_str(), nx, ny, nz, kTeraBlock, dtype))) return fail(tw_tera_last_error()), rc;
    if (p->downsample && (rc = downsample_pass(ctx.c, base, nx, ny, nz, dtype))) return rc;
    if (out_path && out_cap > 0) {
        std::strncpy(out_path, base.c_str(), (size_t)out_cap - 1);
        out_path[out_cap - 1] = '\0';
    }
    return TW_OK;
}

bool is_dir(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

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
    static const char* const suffix[] = {"_affine", "_stps", "_scaling", "_rotation", "_translation"};
    // 2. subject image: a directory of slices takes the tera branch (211-330),
    // a .v3draw stack the in-memory one (155-210).
    if (is_dir(img)) return tera_branch(img, out_dir, p, tar, sub, nsub, suffix[p->kind], out_path, out_cap);
    if (!ends_with(img, ".v3draw"))
        return fail("the subject must be a .v3draw stack or a directory of slices"), TW_E_ARG;
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
