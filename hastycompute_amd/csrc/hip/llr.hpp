// Locally-low-rank SVT kernels (Gram / Jacobi / apply), see llr.hip.
#pragma once

#include "common.hpp"

namespace hasty_hip {

// Casorati addressing of one block: element (r, f) of the block at corner c lives at
//   f * fstride + rowaddr(r) * rstride,  r = ((e*bx + ix)*by + iy)*bz + iz,
//   rowaddr = e*X*Y*Z + (c.x+ix)*Y*Z + (c.y+iy)*Z + (c.z+iz).
struct LlrGeom {
    int64_t F, E, X, Y, Z;
    int bx, by, bz;
    int64_t fstride;
    int64_t rstride;
};

int llr_padded(int64_t F);
size_t llr_workspace_bytes(int64_t nblk, int64_t F, int nsplit);
void llr_check(int64_t F);
void llr_set_max_sweeps(int sweeps);  // Jacobi sweep cap, 1..15 (default 15)
int llr_get_max_sweeps();
void llr_set_eig_threads(int threads);  // eigensolver workgroup: 256, 512 (default) or 768
int llr_get_eig_threads();
void llr_blocks(float2* data, const LlrGeom& g, const int3* corners_dev, int64_t nblk, float thresh, int soft, int rank,
                void* ws, size_t ws_bytes, int nsplit, float* eig_out, hipStream_t s);

}  // namespace hasty_hip
