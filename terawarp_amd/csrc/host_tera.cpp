// host_tera.cpp -- the tera driver's on-disk layout: output blocks and
// slices, and slice-directory input.
//
// The reference's tera-block drivers write each warped block as
//   <base>_block/ZZZZZZ_YYYYYY_XXXXXX.v3draw
// with the block's z, y, x offsets times 10, zero-padded to six digits
// (q_imgwarp_tps_quicksmallmemory.cpp:1239-1241, 1863-1875;
// q_warp_affine_tps.cpp:914-916, 1257, 1545-1547, 1933), and then stitch the
// blocks into one 2-D stack per z,
//   <base>_slice/ZZZZZZZ.raw
// with z times 10 zero-padded to seven digits (Bigwarp.cpp:960-1000,
// written by saveImage_new as a one-slice Vaa3D raw stack).  <base> carries
// the pass's suffix: <out>/<name>_stps, _affine, or <...>_downsample for the
// preview pass.  tw_block_file_write is a tw_write_fn for tw_warp_stream
// that writes the block files; tw_stitch_slices reads them back slice by
// slice (a slice of a block file is contiguous: the file is x-fastest) and
// writes the slices.  A slice directory (the drivers' input,
// read_2Dtif_dir, dir_make.cpp:175-217) is read window by window through
// tw_slice_dir_read.
#include <dirent.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/terawarp.h"
#include "../../include/terawarp_host.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return TW_E_IO;
}

// mkdir -p (the reference shells out to `mkdir -p`, Bigwarp.cpp:35-64).
int make_dirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); i++) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && cur != "/" && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                return fail("cannot create directory " + cur + ": " + strerror(errno));
        }
        if (i < path.size()) cur += path[i];
    }
    return TW_OK;
}

std::string padded(long long v, int width) {
    char b[32];
    snprintf(b, sizeof b, "%0*lld", width, v);
    return b;
}

constexpr int kHeader = 24 + 1 + 2 + 16;  // key, endian byte, datatype, four extents

}  // namespace

struct tw_slice_dir {
    std::vector<std::string> files;  // one slice per z, sorted by name
    int64_t nx = 0, ny = 0;
    int bpv = 1;
    std::vector<int> hl;             // header length of each file
    std::vector<char> swap;          // 16-bit data in the other byte order
};

extern "C" {

int tw_tera_block_path(const char* base, int64_t x0, int64_t y0, int64_t z0, char* out, int cap) {
    if (!base || !out || cap < 1) return TW_E_ARG;
    const std::string p = std::string(base) + "_block/" + padded(z0 * 10, 6) + "_" + padded(y0 * 10, 6) + "_" +
                          padded(x0 * 10, 6) + ".v3draw";
    if ((int)p.size() >= cap) return TW_E_ARG;
    memcpy(out, p.c_str(), p.size() + 1);
    return TW_OK;
}

int tw_tera_slice_path(const char* base, int64_t z, char* out, int cap) {
    if (!base || !out || cap < 1) return TW_E_ARG;
    const std::string p = std::string(base) + "_slice/" + padded(z * 10, 7) + ".raw";
    if ((int)p.size() >= cap) return TW_E_ARG;
    memcpy(out, p.c_str(), p.size() + 1);
    return TW_OK;
}

int tw_block_file_write(void* user, int64_t brick, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                        int64_t nz, const void* data) {
    (void)brick;
    const tw_block_files* f = (const tw_block_files*)user;
    if (!f || !f->base || (f->dtype != TW_U8 && f->dtype != TW_U16)) return TW_E_ARG;
    int rc;
    if ((rc = make_dirs(std::string(f->base) + "_block"))) return rc;
    char path[4096];
    if ((rc = tw_tera_block_path(f->base, x0, y0, z0, path, sizeof path))) return rc;
    const int64_t sz[4] = {nx, ny, nz, 1};
    return tw_v3draw_write(path, data, sz, f->dtype == TW_U8 ? 1 : 2);
}

int tw_stitch_slices(const char* base, int64_t nx, int64_t ny, int64_t nz, int32_t brick, int32_t dtype) {
    if (!base || nx < 1 || ny < 1 || nz < 1 || brick < 1 || (dtype != TW_U8 && dtype != TW_U16)) return TW_E_ARG;
    const int bpv = dtype == TW_U8 ? 1 : 2;
    int rc;
    if ((rc = make_dirs(std::string(base) + "_slice"))) return rc;
    const int64_t nbx = (nx + brick - 1) / brick, nby = (ny + brick - 1) / brick;
    std::vector<uint8_t> slice((size_t)(nx * ny * bpv)), row;
    char path[4096];
    for (int64_t z = 0; z < nz; z++) {
        const int64_t bz = z / brick, z0 = bz * brick;
        for (int64_t by = 0; by < nby; by++)
            for (int64_t bx = 0; bx < nbx; bx++) {
                const int64_t x0 = bx * brick, y0 = by * brick;
                const int64_t bnx = std::min<int64_t>(brick, nx - x0), bny = std::min<int64_t>(brick, ny - y0);
                if ((rc = tw_tera_block_path(base, x0, y0, z0, path, sizeof path))) return rc;
                int64_t sz[4];
                int dt = 0, swap = 0, hl = 0;
                if ((rc = tw_v3draw_header(path, sz, &dt, &swap, &hl))) return fail("missing block " + std::string(path));
                if (sz[0] != bnx || sz[1] != bny || dt != (dtype == TW_U8 ? 1 : 2) || swap)
                    return fail("block " + std::string(path) + " does not match the volume");
                FILE* fp = fopen(path, "rb");
                if (!fp) return fail("cannot open " + std::string(path));
                row.resize((size_t)(bnx * bny * bpv));
                const long off = (long)hl + (long)((z - z0) * bnx * bny * bpv);
                const bool ok = fseek(fp, off, SEEK_SET) == 0 && fread(row.data(), 1, row.size(), fp) == row.size();
                fclose(fp);
                if (!ok) return fail("short read in " + std::string(path));
                for (int64_t y = 0; y < bny; y++)
                    memcpy(slice.data() + (size_t)(((y0 + y) * nx + x0) * bpv), row.data() + (size_t)(y * bnx * bpv),
                           (size_t)(bnx * bpv));
            }
        if ((rc = tw_tera_slice_path(base, z, path, sizeof path))) return rc;
        const int64_t sz[4] = {nx, ny, 1, 1};
        if ((rc = tw_v3draw_write(path, slice.data(), sz, dtype == TW_U8 ? 1 : 2))) return rc;
    }
    return TW_OK;
}

int tw_slice_dir_open(const char* dir, tw_slice_dir** out) {
    if (!dir || !out) return TW_E_ARG;
    *out = nullptr;
    DIR* dh = opendir(dir);
    if (!dh) return fail(std::string("cannot open slice directory ") + dir);
    std::vector<std::string> names;
    while (dirent* e = readdir(dh)) {
        const std::string n = e->d_name;
        if (n != "." && n != ".." && n != "mdata.bin" && n != ".iim.format") names.push_back(n);
    }
    closedir(dh);
    std::sort(names.begin(), names.end());  // byte-wise, as std::list<string>::sort
    if (names.empty()) return fail(std::string("no slices in ") + dir);
    auto* d = new tw_slice_dir;
    d->files.reserve(names.size());
    for (const auto& n : names) d->files.push_back(std::string(dir) + "/" + n);
    for (size_t z = 0; z < d->files.size(); z++) {
        int64_t sz[4];
        int dt = 0, sw = 0, hl = 0;
        if (tw_v3draw_header(d->files[z].c_str(), sz, &dt, &sw, &hl)) {
            delete d;
            return fail("slice " + names[z] + ": not a Vaa3D raw stack");
        }
        if (z == 0) {
            d->nx = sz[0];
            d->ny = sz[1];
            d->bpv = dt;
        }
        if (sz[0] != d->nx || sz[1] != d->ny || sz[2] != 1 || sz[3] != 1 || dt != d->bpv || (dt != 1 && dt != 2)) {
            delete d;
            return fail("slice " + names[z] + ": extent or datatype differs from the first slice (one-channel "
                        "uint8 / uint16 slices of one size)");
        }
        d->hl.push_back(hl);
        d->swap.push_back((char)sw);
    }
    *out = d;
    return TW_OK;
}

int tw_slice_dir_size(const tw_slice_dir* d, int64_t* nx, int64_t* ny, int64_t* nz, int32_t* dtype) {
    if (!d) return TW_E_ARG;
    if (nx) *nx = d->nx;
    if (ny) *ny = d->ny;
    if (nz) *nz = (int64_t)d->files.size();
    if (dtype) *dtype = d->bpv == 1 ? TW_U8 : TW_U16;
    return TW_OK;
}

int tw_slice_dir_read(void* user, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz, void* out) {
    const tw_slice_dir* d = (const tw_slice_dir*)user;
    if (!d || !out || x0 < 0 || y0 < 0 || z0 < 0 || nx < 1 || ny < 1 || nz < 1 || x0 + nx > d->nx ||
        y0 + ny > d->ny || z0 + nz > (int64_t)d->files.size())
        return TW_E_ARG;
    const int bpv = d->bpv;
    uint8_t* o = (uint8_t*)out;
    for (int64_t z = 0; z < nz; z++) {
        const size_t s = (size_t)(z0 + z);
        FILE* fp = fopen(d->files[s].c_str(), "rb");
        if (!fp) return fail("cannot open " + d->files[s]);
        bool ok = true;
        uint8_t* plane = o + (size_t)(z * ny * nx * bpv);
        if (nx == d->nx) {  // whole rows: one read
            ok = fseek(fp, (long)d->hl[s] + (long)(y0 * d->nx * bpv), SEEK_SET) == 0 &&
                 fread(plane, 1, (size_t)(ny * nx * bpv), fp) == (size_t)(ny * nx * bpv);
        } else {
            for (int64_t y = 0; y < ny && ok; y++)
                ok = fseek(fp, (long)d->hl[s] + (long)(((y0 + y) * d->nx + x0) * bpv), SEEK_SET) == 0 &&
                     fread(plane + (size_t)(y * nx * bpv), 1, (size_t)(nx * bpv), fp) == (size_t)(nx * bpv);
        }
        fclose(fp);
        if (!ok) return fail("short read in " + d->files[s]);
        if (d->swap[s] && bpv == 2)
            for (int64_t i = 0; i < nx * ny; i++) std::swap(plane[2 * i], plane[2 * i + 1]);
    }
    return TW_OK;
}

void tw_slice_dir_close(tw_slice_dir* d) { delete d; }

int tw_downsample_factor(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t m = nx > ny ? std::max(nx, nz) : std::max(ny, nz);
    return (int)(m / 600 + (m % 600 >= 300 ? 1 : 0));
}

const char* tw_tera_last_error(void) { return g_err.c_str(); }

}  // extern "C"

static_assert(kHeader == 43, "Vaa3D raw header");
