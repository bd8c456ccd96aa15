// thin3d.h -- host-side interface of the 3D thinning engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bgt {

struct Item3D;
struct Tile3D;

struct Config3D {
    int waves = 8;           // wave64s per workgroup (stacked along z)
    int iters_per_pass = 1;  // K: iterations fused per HBM pass (halo 6K)
};

struct Plan3D {
    int w = 0, h = 0, d = 0;
    int nh = 0;        // 32-bit words per x-line in HBM
    int NH = 0, R = 0; // register words per lane-line and z-rows per lane
    int nbx = 1, nby = 0, nbz = 0, nbricks = 0;
    int iy = 0, iz = 0;
    size_t state_words = 0;
};

int plan3d(int w, int h, int d, const Config3D &cfg, Plan3D &pl);
size_t workspace_bytes3d(const Plan3D &pl);
int setup_item3d(const Plan3D &pl, void *ws, hipStream_t s, Item3D **d_item, unsigned **d_ctr,
                 Item3D *host_item);
int lds_limit();
size_t lds_bytes3d(const Config3D &cfg, int NH);
int rows_for_nh(int nh);
int run_passes3d(const Config3D &cfg, const Plan3D &pl, Item3D *d_item, unsigned *d_ctr,
                 const int *d_state, hipStream_t s, int *iterations);

int thin3d_device_u8(const uint8_t *d_in, int w, int h, int d, uint8_t *d_out, hipStream_t s,
                     int *iterations, const Config3D &cfg);
int thin3d_host(int dtype, int w, int h, int d, const void *in, void *out, const Config3D &cfg);
int threshold_int_positive(int dtype, const void *in, size_t n, uint8_t *mask);

const Config3D &default_config3d();

}  // namespace bgt
