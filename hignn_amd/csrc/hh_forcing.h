// Launch descriptors of the forcing and integrator kernels (hh_forcing.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hh {

struct Lim {
  float lo[3], hi[3];
};

struct ForcingLaunch {
  const float *x;  // (n,3) positions
  float *f;        // (n,3) forces (output)
  int64_t n;
  float b[3];  // body force
  // filament chains: particles [0, n_fil * n_chain) in chains of n_chain
  int64_t n_fil, n_chain;
  float rest, k_t;
  double k_b;
  // Morse bonds between consecutive beads of bond chains: particles
  // [0, bond_fil * bond_chain) in chains of bond_chain; U(r) = D (1 -
  // exp(-alpha (r - r0)))^2 per bond
  int64_t bond_fil, bond_chain;
  double bond_D, bond_alpha, bond_r0;
  // Morse neighbours within eps
  int morse, periodic;
  float eps;
  float margin;  // periodic image margin (the two-body radius)
  double eps2;   // (double)(eps * eps), the float square of the radius
  double inv_h;  // 1 / cell size; cell size = eps (1 + 1e-6)
  float lo[3], hi[3];
  // sources (periodic images or the particles themselves)
  int64_t src_cap;
  float4 *src;
  int32_t *src_tgt;
  int4 *src_cell;
  uint32_t *cnt, *off;  // image counts and offsets (n + 1); off[n] = number of sources
  uint32_t *nsrc;       // number of sources (device; off + n)
  // hashed cell table
  uint32_t tsize, tmask;
  uint32_t *bcnt, *bstart;  // tsize, tsize + 1
  uint32_t *bsum;           // scan partials
  int32_t *sidx;            // sources sorted by bucket
  float4 *spos;
  int4 *scell;
};

// Stage positions x = x0 + sum_k coef[row*8+k] K_k (left to right), then the
// periodic image offset: xs (optional) receives x before the offset, xin after.
struct StageLaunch {
  const float *x0;
  const float *k[8];
  const float *coef;  // device
  int row, nterms;
  float *xs, *xin;
  int64_t n;
  int periodic;
  float lo[3], hi[3];
};

struct ErrLaunch {
  const float *k[7];
  float dc[7];
  const float *fdt;  // device
  double *part;
  int64_t n;
};

// Exclusive scan of n 32-bit counts into out[0..n] (out[n] = total); bsum
// holds scan_partials_needed(n) entries.
hipError_t exclusive_scan(const uint32_t *in, int64_t n, uint32_t *out, uint32_t *bsum,
                          hipStream_t st);
hipError_t launch_forcing(const ForcingLaunch &L, hipStream_t st);
// Sources and hashed cell table of L (the first half of launch_forcing).
hipError_t launch_cell_list(const ForcingLaunch &L, hipStream_t st);
// Neighbour lists within L.eps: counts, then off = exclusive scan (n + 1
// entries, off[n] = total); then the lists, each sorted ascending.
hipError_t launch_neighbour_count(const ForcingLaunch &L, uint32_t *cnt, uint32_t *off,
                                  uint32_t *bsum, hipStream_t st);
hipError_t launch_neighbour_fill(const ForcingLaunch &L, const uint32_t *off, int32_t *idx,
                                 hipStream_t st);
int64_t scan_partials_needed(int64_t n);
hipError_t launch_stage(const StageLaunch &S, hipStream_t st);
hipError_t launch_euler_update(float *x, const float *v, float *off, int64_t n, const float *fdt,
                               int periodic, Lim lim, hipStream_t st);
hipError_t launch_add(float *x, const float *off, int64_t n3, hipStream_t st);
// *p = v in stream order (no pinned staging buffer left in flight).
hipError_t launch_set_f32(float *p, float v, hipStream_t st);
int64_t err_partials_needed(int64_t n);
hipError_t launch_dopri_err(const ErrLaunch &E, double *out, hipStream_t st);

}  // namespace hh
