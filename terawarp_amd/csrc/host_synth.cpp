// host_synth.cpp -- procedural subject volumes and brick checksums for the
// streaming driver (tw_warp_stream).
//
// The reference reads its tera-scale input from slice or block files
// (q_imgwarp_tps_quicksmallmemory.cpp:1161-1994, Bigwarp.cpp:229-573); a
// volume larger than host memory is stood in for here by a field that is a
// pure function of the seed and the global voxel coordinate, so any window,
// filled by any thread, holds the same values as every overlapping window.
//
// Field: uniform noise from a 32-bit hash of (seed, x, y, z) plus Gaussian
// blobs.  The volume is cut into cells of `cell` voxels; each cell holds
// `blobs_per_cell` blobs (centre uniform in the cell, sigma and amplitude
// uniform in their ranges) drawn from a generator seeded by (seed, cell).  A
// voxel sums the noise and then every blob whose 3-sigma box holds it, in
// (cell z, y, x, blob) order, in FP32; the sum is clamped to the dtype's
// range and rounded half up.
// The loops below are written to vectorise; -O2 of this compiler does not.
#pragma GCC optimize("O3")

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/terawarp.h"

namespace {

// 32-bit finaliser (xorshift-multiply, two rounds): vectorises over x.
inline uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// Noise of one row: x0 .. x0 + n - 1 at row key `row`.
__attribute__((target_clones("avx2", "default"))) void noise_row(float* a, int64_t n, int64_t x0, uint32_t row,
                                                                 float amp) {
    for (int64_t x = 0; x < n; x++) {
        const uint32_t h = mix32(row ^ ((uint32_t)(x0 + x) * 0x9E3779B1u));
        a[x] = amp * ((float)(int32_t)(h >> 16) * (1.0f / 32768.0f) - 1.0f);
    }
}

// o = clamp(round half up(a), 0, vmax): rounding first and clamping the
// integer gives the same values as the other order and keeps the loop free
// of branches.
template <typename T>
__attribute__((target_clones("avx2", "default"))) void store_row(const float* a, T* o, size_t n, int32_t vmax) {
    for (size_t e = 0; e < n; e++) {
        int32_t i = (int32_t)(a[e] + 0.5f);  // |a| stays far below 2^31
        i = i < 0 ? 0 : i;
        i = i > vmax ? vmax : i;
        o[e] = (T)i;
    }
}

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Blob {
    float cx, cy, cz, s, a;
    int64_t x0, x1, y0, y1, z0, z1;  // 3-sigma box (inclusive lower, exclusive upper)
};

// The blobs of cell (i, j, k), in a fixed order.
void cell_blobs(const tw_synth& s, int64_t i, int64_t j, int64_t k, std::vector<Blob>& out) {
    uint64_t st = mix64(s.seed ^ mix64(((uint64_t)k * 0x100000001B3ull + (uint64_t)j) * 0x100000001B3ull +
                                       (uint64_t)i));
    auto uni = [&]() {
        st = mix64(st);
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int b = 0; b < s.blobs_per_cell; b++) {
        Blob B;
        B.cx = (float)((i + uni()) * s.cell);
        B.cy = (float)((j + uni()) * s.cell);
        B.cz = (float)((k + uni()) * s.cell);
        B.s = (float)(s.sigma_lo + (s.sigma_hi - s.sigma_lo) * uni());
        B.a = (float)(s.amp_lo + (s.amp_hi - s.amp_lo) * uni());
        const int64_t r = (int64_t)std::ceil(3.0f * B.s);
        B.x0 = (int64_t)std::floor(B.cx) - r; B.x1 = (int64_t)std::floor(B.cx) + r + 1;
        B.y0 = (int64_t)std::floor(B.cy) - r; B.y1 = (int64_t)std::floor(B.cy) + r + 1;
        B.z0 = (int64_t)std::floor(B.cz) - r; B.z1 = (int64_t)std::floor(B.cz) + r + 1;
        out.push_back(B);
    }
}

template <typename T>
void fill(const tw_synth& s, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz, T* out) {
    const float vmax = sizeof(T) == 1 ? 255.f : 65535.f;  // the sum is clamped to [0, vmax]
    // Blobs that can reach the box, in (cell z, y, x, blob) order.
    const int64_t reach = (int64_t)std::ceil(3.0f * s.sigma_hi) + 1;
    auto cells = [&](int64_t a0, int64_t n, int64_t lim) {
        const int64_t c0 = std::max<int64_t>(0, (a0 - reach) / s.cell);
        const int64_t c1 = std::min<int64_t>((lim + s.cell - 1) / s.cell - 1, (a0 + n + reach) / s.cell);
        return std::make_pair(c0, c1);
    };
    const auto cx = cells(x0, nx, s.nx), cy = cells(y0, ny, s.ny), cz = cells(z0, nz, s.nz);
    std::vector<Blob> blobs;
    for (int64_t k = cz.first; k <= cz.second; k++)
        for (int64_t j = cy.first; j <= cy.second; j++)
            for (int64_t i = cx.first; i <= cx.second; i++) cell_blobs(s, i, j, k, blobs);
    std::vector<Blob> hit;
    for (const Blob& b : blobs)
        if (b.x1 > x0 && b.x0 < x0 + nx && b.y1 > y0 && b.y0 < y0 + ny && b.z1 > z0 && b.z0 < z0 + nz)
            hit.push_back(b);
    std::vector<float> acc((size_t)nx * ny), gx, gy;
    for (int64_t z = 0; z < nz; z++) {
        const int64_t Z = z0 + z;
        // Noise: a row key from (seed, y, z), mixed with x per voxel.
        for (int64_t y = 0; y < ny; y++) {
            const uint32_t row = (uint32_t)mix64(s.seed ^ ((uint64_t)Z << 32) ^ (uint64_t)(y0 + y));
            noise_row(acc.data() + (size_t)y * nx, nx, x0, row, s.noise);
        }
        // Blobs through this slice.
        for (const Blob& b : hit) {
            if (Z < b.z0 || Z >= b.z1) continue;
            const float inv = 1.0f / (2.0f * b.s * b.s);
            const float dz = (float)Z - b.cz;
            const float wz = b.a * std::exp(-dz * dz * inv);
            const int64_t xa = std::max(b.x0, x0), xb = std::min(b.x1, x0 + nx);
            const int64_t ya = std::max(b.y0, y0), yb = std::min(b.y1, y0 + ny);
            gx.resize((size_t)(xb - xa));
            gy.resize((size_t)(yb - ya));
            for (int64_t x = xa; x < xb; x++) {
                const float d = (float)x - b.cx;
                gx[(size_t)(x - xa)] = std::exp(-d * d * inv);
            }
            for (int64_t y = ya; y < yb; y++) {
                const float d = (float)y - b.cy;
                gy[(size_t)(y - ya)] = wz * std::exp(-d * d * inv);
            }
            for (int64_t y = ya; y < yb; y++) {
                float* a = acc.data() + (size_t)(y - y0) * nx + (xa - x0);
                const float w = gy[(size_t)(y - ya)];
                for (int64_t x = 0; x < xb - xa; x++) a[x] += w * gx[(size_t)x];
            }
        }
        store_row(acc.data(), out + (size_t)z * nx * ny, acc.size(), (int32_t)vmax);
    }
}

}  // namespace

extern "C" int tw_synth_fill(const tw_synth* s, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                             int64_t nz, void* out) {
    if (!s || !out || nx < 0 || ny < 0 || nz < 0 || s->cell < 1 || s->blobs_per_cell < 0) return TW_E_ARG;
    if (x0 < 0 || y0 < 0 || z0 < 0 || x0 + nx > s->nx || y0 + ny > s->ny || z0 + nz > s->nz) return TW_E_ARG;
    if (s->dtype == TW_U8) fill(*s, x0, y0, z0, nx, ny, nz, (uint8_t*)out);
    else if (s->dtype == TW_U16) fill(*s, x0, y0, z0, nx, ny, nz, (uint16_t*)out);
    else return TW_E_ARG;
    return TW_OK;
}

extern "C" int tw_synth_read(void* user, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz,
                             void* out) {
    return tw_synth_fill((const tw_synth*)user, x0, y0, z0, nx, ny, nz, out);
}

// Sum over the 64-bit little-endian words w_i of the data (the last one
// zero-padded) of w_i * (2i + 1), modulo 2^64, xor the byte count: every
// word and its position count.
namespace {

// sum of w_i * (2i + 1) over the whole words i in [i0, i1), four chains.
uint64_t word_sum(const uint8_t* p, int64_t i0, int64_t i1) {
    uint64_t h[4] = {0, 0, 0, 0};
    int64_t i = i0;
    for (; i + 4 <= i1; i += 4)
        for (int k = 0; k < 4; k++) {
            uint64_t w;
            memcpy(&w, p + 8 * (i + k), 8);
            h[k] += w * (2 * (uint64_t)(i + k) + 1);
        }
    for (; i < i1; i++) {
        uint64_t w;
        memcpy(&w, p + 8 * i, 8);
        h[0] += w * (2 * (uint64_t)i + 1);
    }
    return h[0] + h[1] + h[2] + h[3];
}

}  // namespace

extern "C" uint64_t tw_checksum64(const void* data, int64_t bytes) {
    const uint8_t* p = (const uint8_t*)data;
    const int64_t nw = bytes / 8;
    // Large buffers in parallel parts (the sum is exact modulo 2^64 in any order).
    const int parts = nw >= (1 << 22) ? 4 : 1;
    uint64_t h = 0;
    if (parts > 1) {
        uint64_t part[4] = {0, 0, 0, 0};
        std::vector<std::thread> th;
        for (int t = 0; t < parts; t++)
            th.emplace_back([&, t] { part[t] = word_sum(p, nw * t / parts, nw * (t + 1) / parts); });
        for (auto& x : th) x.join();
        for (int t = 0; t < parts; t++) h += part[t];
    } else {
        h = word_sum(p, 0, nw);
    }
    if (bytes % 8) {
        uint64_t w = 0;
        memcpy(&w, p + 8 * nw, (size_t)(bytes % 8));
        h += w * (2 * (uint64_t)nw + 1);
    }
    return h ^ (uint64_t)bytes;
}

extern "C" int tw_checksum_write(void* user, int64_t brick, int64_t x0, int64_t y0, int64_t z0, int64_t nx,
                                 int64_t ny, int64_t nz, const void* data) {
    (void)x0; (void)y0; (void)z0;
    tw_checksum_sink* k = (tw_checksum_sink*)user;
    if (!k || brick < 0 || brick >= k->nbricks) return TW_E_ARG;
    const int64_t bytes = nx * ny * nz * (int64_t)k->bytes_per_voxel;
    if (k->sums) k->sums[brick] = tw_checksum64(data, bytes);
    for (int i = 0; i < k->nkeep; i++)
        if (k->keep_ids[i] == brick && k->keep_ptrs[i]) memcpy(k->keep_ptrs[i], data, (size_t)bytes);
    return TW_OK;
}
