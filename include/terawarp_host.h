/*
 * terawarp_host.h -- host-side mirror of the reference's plugin surface and
 * file I/O, in libterawarp.so above the compute ABI of terawarp.h.
 *
 * The reference is a Vaa3D plugin (plugin_Bigimagewarp.{h,cpp}); Vaa3D and
 * Qt are not linked here, so the plugin's entry points are exported with
 * plain C types: the QStringList / V3DPluginArgList vectors become arrays
 * of C strings.  A Vaa3D build wraps these in its V3DPluginInterface2_1
 * class (INTEGRATION.md shows the shim).
 */
#ifndef TERAWARP_HOST_H
#define TERAWARP_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Vaa3D raw stack (.v3draw) --------------------------------------------
 * Header: "raw_image_stack_by_hpeng" (24 bytes), endian byte 'L'/'B', int16
 * datatype (1 = uint8, 2 = uint16, 4 = float32), four int32 extents
 * x, y, z, c; then the data, x fastest (RawFmtMngr.cpp:216-330 reads it,
 * Bigwarp.cpp:660-751 writes it). */

/* Read the header: sz[4] = {x, y, z, c}, *datatype in {1, 2, 4}, *swap = 1
 * when the file's endianness differs from the host's.  Returns TW_OK or
 * TW_E_IO (message via tw_io_last_error). */
int tw_v3draw_header(const char* path, int64_t* sz, int* datatype, int* swap, int* header_len);
/* Read the whole stack into buf (capacity bytes), byte-swapping 16-bit data
 * from the other endianness. */
int tw_v3draw_read(const char* path, void* buf, int64_t capacity);
/* Write a stack in the host's endianness. */
int tw_v3draw_write(const char* path, const void* data, const int64_t* sz, int datatype);

/* ---- Vaa3D markers (.marker) -----------------------------------------------
 * CSV lines "x,y,z,radius,shape,name,comment,r,g,b"; '#' starts a comment
 * line.  Coordinates are used as stored (no 1-based shift), as at
 * plugin_Bigimagewarp.cpp:145-146.  Reads at most max_n points into
 * xyz (3 floats each); returns the number of markers in the file (which may
 * exceed max_n) or a negative status. */
int tw_markers_read(const char* path, float* xyz, int max_n);
int tw_markers_write(const char* path, const float* xyz, int n);

const char* tw_io_last_error(void);

/* ---- tera output layout ---------------------------------------------------------
 * The tera-block driver's files (q_imgwarp_tps_quicksmallmemory.cpp:1239-1241,
 * 1863-1875; Bigwarp.cpp:960-1000): one Vaa3D raw stack per output block,
 *   <base>_block/ZZZZZZ_YYYYYY_XXXXXX.v3draw  (offsets x 10, 6 digits),
 * and one single-slice stack per z,
 *   <base>_slice/ZZZZZZZ.raw                   (z x 10, 7 digits),
 * where <base> carries the pass's suffix (<out>/<name>_stps, _affine,
 * <...>_downsample; q_warp_affine_tps.cpp:914-916, 1545-1547). */
typedef struct {
    const char* base;
    int32_t     dtype;  /* TW_U8 / TW_U16 */
    int32_t     pad;
} tw_block_files;
int tw_tera_block_path(const char* base, int64_t x0, int64_t y0, int64_t z0, char* out, int cap);
int tw_tera_slice_path(const char* base, int64_t z, char* out, int cap);
/* tw_write_fn (include/terawarp.h) writing each brick as its block file
 * (user = const tw_block_files*). */
int tw_block_file_write(void* user, int64_t brick, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                        int64_t nz, const void* data);
/* Read the block files of a volume of nx x ny x nz in bricks of `brick` and
 * write its slices. */
int tw_stitch_slices(const char* base, int64_t nx, int64_t ny, int64_t nz, int32_t brick, int32_t dtype);
/* ---- slice-directory input ------------------------------------------------------
 * A volume stored as one single-slice Vaa3D raw stack per z in a directory:
 * every entry except ".", "..", "mdata.bin" and ".iim.format", sorted by
 * name, is one slice (read_2Dtif_dir, dir_make.cpp:175-217; the slices'
 * header read as loadMetadata_cyf, RawFmtMngr.cpp:332-446).  All slices must
 * share the first one's extent and datatype (uint8 / uint16). */
typedef struct tw_slice_dir tw_slice_dir;
int  tw_slice_dir_open(const char* dir, tw_slice_dir** out);
int  tw_slice_dir_size(const tw_slice_dir* d, int64_t* nx, int64_t* ny, int64_t* nz, int32_t* dtype);
/* tw_read_fn (include/terawarp.h) over an open slice directory (user =
 * tw_slice_dir*); safe to call from several threads at once. */
int  tw_slice_dir_read(void* user, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz,
                       void* out);
void tw_slice_dir_close(tw_slice_dir* d);

/* Downsample factor of the preview pass (q_warp_affine_tps.cpp:1477-1479):
 * with m the largest extent, m / 600 rounded half up; 0 when m < 300 (the
 * reference then divides by zero). */
int  tw_downsample_factor(int64_t nx, int64_t ny, int64_t nz);

const char* tw_tera_last_error(void);

/* ---- plugin surface ----------------------------------------------------------
 * plugin_Bigimagewarp.cpp:26-38 (menulist / funclist), 73-107 (dofunc),
 * 123-336 (bigimagewarp). */

/* Number of entries and entry i of the menu / function lists. */
int         tw_plugin_menu_count(void);
const char* tw_plugin_menu_item(int i);
int         tw_plugin_func_count(void);
const char* tw_plugin_func_item(int i);
float       tw_plugin_version(void);

/* Parsed -p parameter vector.  Two forms are accepted:
 *  - the README form, 10 values (README.md:85-100):
 *      warp_mode device_mode resize_mode deformation_mode
 *      para_x para_y para_z newsize_x newsize_y newsize_z
 *    warp_mode 0 (non-rigid): deformation 0 affine, 1 TPS, 2 scaling;
 *    warp_mode 1 (rigid):     deformation 0 rotation, 1 translation;
 *  - the reference code's form, 7 values of which 3 are read
 *    (plugin_Bigimagewarp.cpp:78, 91-93): program_mode (1 affine, 0 TPS)
 *    device_mode downsample_mode x x x x.
 * device_mode 0 asks for the reference CPU path's results: the GPU runs the
 * FP64 reference-order form (TW_EXACT); device_mode 1 runs the plain GPU
 * form; device_mode 2 the mixed-precision form. */
typedef struct {
    int     kind;        /* 0 affine (landmarks), 1 TPS, 2 scaling, 3 rotation, 4 translation */
    int     device_mode; /* 0, 1, 2 as above */
    int     resize;      /* 1: output extent newsize[] */
    int     downsample;  /* legacy form's downsample_mode (tera-scale path only) */
    double  para[3];
    int64_t newsize[3];
    int     legacy;      /* 1 when parsed from the 7-value form */
} tw_warp_params;

int tw_plugin_parse_params(const char* const* paras, int n, tw_warp_params* out);

/* dofunc("bigimagewarp", input, output): inputs = {image, marker_sub,
 * marker_tar}, paras = the -p vector, outputs = {output folder}.  Returns 1
 * on success, 0 on failure (message via tw_io_last_error), like the
 * reference's bool.  func_name "help" prints the usage. */
int tw_plugin_dofunc(const char* func_name, const char* const* inputs, int n_inputs, const char* const* paras,
                     int n_paras, const char* const* outputs, int n_outputs);

/* The warp behind dofunc for an in-memory .v3draw subject: reads the image
 * and markers, runs the transform on the GPU and writes
 * <out_dir>/<basename><suffix>.v3draw, where suffix is _affine, _stps,
 * _scaling, _rotation or _translation.  The output path is copied to
 * out_path (capacity out_cap) when non-null. */
int tw_bigimagewarp(const char* img_path, const char* marker_sub, const char* marker_tar, const char* out_dir,
                    const tw_warp_params* params, char* out_path, int out_cap);

#ifdef __cplusplus
}
#endif

#endif /* TERAWARP_HOST_H */
