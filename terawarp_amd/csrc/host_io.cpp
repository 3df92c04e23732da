// host_io.cpp -- Vaa3D .v3draw stacks and .marker files.
//
// Formats as read by loadMetadata (RawFmtMngr.cpp:216-330) and written by
// saveStack2Raw_cyf (Bigwarp.cpp:660-751); the marker CSV is Vaa3D's
// (readMarker_file, called at plugin_Bigimagewarp.cpp:145-146).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/terawarp.h"
#include "../../include/terawarp_host.h"

namespace {

thread_local std::string g_err;

int io_fail(const std::string& msg) {
    g_err = msg;
    return TW_E_IO;
}

const char kKey[] = "raw_image_stack_by_hpeng";
constexpr int kKeyLen = 24;

char host_endian() {
    const uint16_t v = 1;
    return *(const uint8_t*)&v == 1 ? 'L' : 'B';
}

void swap2(void* p) {
    uint8_t* b = (uint8_t*)p;
    std::swap(b[0], b[1]);
}

void swap4(void* p) {
    uint8_t* b = (uint8_t*)p;
    std::swap(b[0], b[3]);
    std::swap(b[1], b[2]);
}

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) { f = std::fopen(path, mode); }
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

namespace tw_host {
void set_error(const std::string& m) { g_err = m; }
}

extern "C" {

const char* tw_io_last_error(void) { return g_err.c_str(); }

int tw_v3draw_header(const char* path, int64_t* sz, int* datatype, int* swap, int* header_len) {
    if (!path || !sz || !datatype) return io_fail("null argument");
    File fh(path, "rb");
    if (!fh.f) return io_fail(std::string("Fail to open file for reading: ") + path);
    std::fseek(fh.f, 0, SEEK_END);
    const long fsize = std::ftell(fh.f);
    std::rewind(fh.f);
    if (fsize < kKeyLen + 2 + 4 * 4 + 1) return io_fail("file too small to contain a v3draw header");
    char key[kKeyLen + 1] = {0};
    if (std::fread(key, 1, kKeyLen, fh.f) != (size_t)kKeyLen || std::strcmp(key, kKey) != 0)
        return io_fail("Unrecognized file format");
    char endian = 0;
    if (std::fread(&endian, 1, 1, fh.f) != 1 || (endian != 'B' && endian != 'L'))
        return io_fail("only big- or little-endian v3draw files are supported");
    const int sw = endian != host_endian();
    int16_t dcode = 0;
    if (std::fread(&dcode, 2, 1, fh.f) != 1) return io_fail("truncated header");
    if (sw) swap2(&dcode);
    if (dcode != 1 && dcode != 2 && dcode != 4) return io_fail("Unrecognized data type code " + std::to_string(dcode));
    int32_t dims[4];
    if (std::fread(dims, 4, 4, fh.f) != 4) return io_fail("truncated header");
    for (int i = 0; i < 4; i++) {
        if (sw) swap4(&dims[i]);
        if (dims[i] <= 0) return io_fail("non-positive extent in header");
        sz[i] = dims[i];
    }
    const long hl = std::ftell(fh.f);
    const int64_t need = hl + sz[0] * sz[1] * sz[2] * sz[3] * (int64_t)dcode;
    if (fsize < need) return io_fail("file shorter than its header declares");
    *datatype = dcode;
    if (swap) *swap = sw;
    if (header_len) *header_len = (int)hl;
    return TW_OK;
}

int tw_v3draw_read(const char* path, void* buf, int64_t capacity) {
    int64_t sz[4];
    int dt = 0, sw = 0, hl = 0;
    int rc = tw_v3draw_header(path, sz, &dt, &sw, &hl);
    if (rc) return rc;
    const int64_t n = sz[0] * sz[1] * sz[2] * sz[3];
    if (!buf || capacity < n * dt) return io_fail("buffer too small for the stack");
    File fh(path, "rb");
    if (!fh.f) return io_fail(std::string("Fail to open file for reading: ") + path);
    std::fseek(fh.f, hl, SEEK_SET);
    if ((int64_t)std::fread(buf, dt, (size_t)n, fh.f) != n) return io_fail("short read of the image data");
    if (sw && dt == 2)
        for (int64_t i = 0; i < n; i++) swap2((uint8_t*)buf + 2 * i);
    if (sw && dt == 4)
        for (int64_t i = 0; i < n; i++) swap4((uint8_t*)buf + 4 * i);
    return TW_OK;
}

int tw_v3draw_write(const char* path, const void* data, const int64_t* sz, int datatype) {
    if (!path || !data || !sz) return io_fail("null argument");
    if (datatype != 1 && datatype != 2 && datatype != 4) return io_fail("datatype must be 1, 2 or 4");
    File fh(path, "wb");
    if (!fh.f) return io_fail(std::string("Fail to open file for writing: ") + path + " (" + std::strerror(errno) + ")");
    const char e = host_endian();
    const int16_t dcode = (int16_t)datatype;
    int32_t dims[4];
    for (int i = 0; i < 4; i++) {
        if (sz[i] <= 0 || sz[i] > 0x7fffffff) return io_fail("extent out of range");
        dims[i] = (int32_t)sz[i];
    }
    const int64_t n = sz[0] * sz[1] * sz[2] * sz[3];
    if (std::fwrite(kKey, 1, kKeyLen, fh.f) != (size_t)kKeyLen || std::fwrite(&e, 1, 1, fh.f) != 1 ||
        std::fwrite(&dcode, 2, 1, fh.f) != 1 || std::fwrite(dims, 4, 4, fh.f) != 4 ||
        (int64_t)std::fwrite(data, datatype, (size_t)n, fh.f) != n)
        return io_fail("write error");
    return TW_OK;
}

int tw_markers_read(const char* path, float* xyz, int max_n) {
    if (!path) return io_fail("null path");
    std::ifstream in(path);
    if (!in) return io_fail(std::string("cannot open marker file ") + path);
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '#') continue;
        std::stringstream ss(line);
        std::string f[3];
        int k = 0;
        for (; k < 3 && std::getline(ss, f[k], ','); k++) {
        }
        if (k < 3) continue;  // fewer than three fields: not a marker line
        float v[3];
        bool ok = true;
        for (int a = 0; a < 3; a++) {
            char* end = nullptr;
            v[a] = std::strtof(f[a].c_str(), &end);
            if (end == f[a].c_str()) ok = false;
        }
        if (!ok) return io_fail("malformed marker line: " + line);
        if (xyz && n < max_n)
            for (int a = 0; a < 3; a++) xyz[3 * n + a] = v[a];
        n++;
    }
    return n;
}

int tw_markers_write(const char* path, const float* xyz, int n) {
    if (!path || (!xyz && n > 0) || n < 0) return io_fail("bad arguments");
    File fh(path, "w");
    if (!fh.f) return io_fail(std::string("cannot write ") + path);
    std::fprintf(fh.f, "##x,y,z,radius,shape,name,comment, color_r,color_g,color_b\n");
    for (int i = 0; i < n; i++)
        std::fprintf(fh.f, "%.9g,%.9g,%.9g,0,1,,,255,0,0\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return TW_OK;
}

}  // extern "C"
