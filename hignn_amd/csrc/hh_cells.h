// Device buffers of a hashed cell list (sources, periodic images, buckets),
// shared by the forcing program and the neighbour lists.
#pragma once

#include <algorithm>
#include <array>

#include "hh_forcing.h"
#include "hh_host.h"

namespace hh {

struct CellBufs {
  DBuf src, src_tgt, src_cell, cnt, off, bcnt, bstart, bsum, sidx, spos, scell;

  // Fills the cell-list part of L for n particles (L.eps, L.periodic,
  // L.lo/hi, L.margin already set), reserving the buffers.
  void setup(ForcingLaunch &L, int64_t n) {
    L.eps2 = (double)(L.eps * L.eps);
    L.inv_h = L.eps > 0 ? 1.0 / ((double)L.eps * (1.0 + 1e-6)) : 0.0;
    if (L.periodic)
      for (int j = 0; j < 3; j++)
        if (!(L.hi[j] - L.lo[j] > 2.0f * L.margin))
          throw HHError(HH_ERR_ARG, "periodic domain must be wider than twice the radius");
    const int64_t cap = L.periodic ? 8 * n : n;
    if (cap >= (int64_t)1 << 30) throw HHError(HH_ERR_ARG, "too many particles for the cell list");
    uint32_t T = 1024;
    while ((int64_t)T < 2 * cap) T <<= 1;
    src.reserve(16 * (size_t)cap);
    src_tgt.reserve(4 * (size_t)cap);
    src_cell.reserve(16 * (size_t)cap);
    sidx.reserve(4 * (size_t)cap);
    spos.reserve(16 * (size_t)cap);
    scell.reserve(16 * (size_t)cap);
    cnt.reserve(4 * (size_t)std::max<int64_t>(n, 1));
    off.reserve(4 * (size_t)(n + 1));
    bcnt.reserve(4 * (size_t)T);
    bstart.reserve(4 * ((size_t)T + 1));
    bsum.reserve(4 * (size_t)scan_partials_needed(std::max<int64_t>(n, T)));
    L.src_cap = cap;
    L.src = src.as<float4>();
    L.src_tgt = src_tgt.as<int32_t>();
    L.src_cell = src_cell.as<int4>();
    L.cnt = cnt.as<uint32_t>();
    L.off = off.as<uint32_t>();
    L.nsrc = off.as<uint32_t>() + n;
    L.tsize = T;
    L.tmask = T - 1;
    L.bcnt = bcnt.as<uint32_t>();
    L.bstart = bstart.as<uint32_t>();
    L.bsum = bsum.as<uint32_t>();
    L.sidx = sidx.as<int32_t>();
    L.spos = spos.as<float4>();
    L.scell = scell.as<int4>();
  }

  std::array<void *, 11> pointers() const {
    return {src.p, src_tgt.p, src_cell.p, cnt.p, off.p, bcnt.p, bstart.p, bsum.p, sidx.p, spos.p,
            scell.p};
  }

  void release() {
    for (DBuf *d : {&src, &src_tgt, &src_cell, &cnt, &off, &bcnt, &bstart, &bsum, &sidx, &spos,
                    &scell})
      d->release();
  }
};

}  // namespace hh
