// Neighbour maintenance for the packed 64-atom cluster list (gfx950).
//
// Reference behaviour (what is kept):
//   * displacement check |x - x_ref|^2 > (r_skin/2)^2 sets bit 0 of a status
//     word (kernels/neighbor.cu:50-75);
//   * the list holds, per i-block of consecutive sorted atoms, "entries" of
//     individual j-atoms, each j-atom within r_search of at least one i-atom,
//     j-blocks >= i-block only (kernels/neighbor.cu:416-631);
//   * capacity overflow / high-water bits in the same status word
//     (kernels/neighbor.cu:643-671).
// MI355X design (what changes):
//   * blocks are 64 atoms = one wave64; entries hold 64 j-atoms;
//   * atoms are ordered along a Hilbert curve over a fixed voxel grid
//     (voxel = fractional-coordinate cell, so the order survives box
//     rescaling) instead of by particle id inside r_search cells, which
//     makes a block spatially compact;
//   * each i-block owns a fixed-capacity entry region, written in a fixed
//     order (no atomics, no scatter pass), so the list is deterministic;
//   * every kernel here is gated on the status word on the device, so the
//     whole rebuild lives inside a captured step and needs no host sync.
#include "hd_device.h"
#include "hd_kernels.h"


__device__ __forceinline__ bool nb_rebuild_due(const unsigned* status) {
  return (*(volatile const unsigned*)status & (NB_STATUS_REBUILD | NB_STATUS_FORCE)) != 0u;
}

// Displacement check against the positions of the last rebuild.  A box
// change since the rebuild (barostat) moves every minimum-image pair
// vector by at most sqrt(3) * max|dL|, which is added to each atom's
// displacement so that |x - x_ref| + sqrt(3) max|dL| <= r_skin/2 still
// bounds every pair-distance change by r_skin.
__global__ void k_disp_check(const float4* __restrict__ posq, const float4* __restrict__ ref,
                             const float* __restrict__ lattice,
                             const float* __restrict__ ref_lattice, float half_skin,
                             unsigned* status, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  bool over = false;
  if (i < n) {
    HdLattice L = hd_load_lattice(lattice);
    float dl = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) dl = fmaxf(dl, fabsf(lattice[k] - ref_lattice[k]));
    float4 p = posq[i], r = ref[i];
    float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
    hd_min_image(dx, dy, dz, L);
    over = sqrtf(dx * dx + dy * dy + dz * dz) + 1.7320508f * dl > half_skin;
  }
  if (__any(over) && (threadIdx.x & 63) == 0) atomicOr(status, NB_STATUS_REBUILD);
}

// Curve rank of the fractional voxel that holds the atom (wrapped into the
// primary cell).
__device__ __forceinline__ unsigned nb_voxel_key(float4 p, const HdLattice& L,
                                                 const unsigned* __restrict__ voxel_rank,
                                                 unsigned nva, unsigned nvb, unsigned nvc) {
  float x = p.x, y = p.y, z = p.z;
  int ia, ib, ic;
  hd_wrap_with_image(x, y, z, ia, ib, ic, L);
  float sa, sb, sc;
  hd_cart_to_frac(x, y, z, L, sa, sb, sc);
  int va = (int)floorf((sa + 0.5f) * (float)nva);
  int vb = (int)floorf((sb + 0.5f) * (float)nvb);
  int vc = (int)floorf((sc + 0.5f) * (float)nvc);
  va = min(max(va, 0), (int)nva - 1);
  vb = min(max(vb, 0), (int)nvb - 1);
  vc = min(max(vc, 0), (int)nvc - 1);
  return voxel_rank[((unsigned)va * nvb + (unsigned)vb) * nvc + (unsigned)vc];
}

// Voxel key = curve rank of the atom's fractional voxel; histogram.
// Also records the positions and lattice of this rebuild, against which the
// displacement check measures (positions do not change within a force
// evaluation, so the copy can be taken here, with the keys).
__global__ void k_voxel_keys(const float4* __restrict__ posq, const float* __restrict__ lattice,
                             const unsigned* __restrict__ voxel_rank, unsigned nva, unsigned nvb,
                             unsigned nvc, unsigned* __restrict__ atom_key,
                             unsigned* __restrict__ voxel_count, float4* __restrict__ ref,
                             float* __restrict__ ref_lattice, const unsigned* status,
                             unsigned n) {
  if (!nb_rebuild_due(status)) return;
  HdLattice L = hd_load_lattice(lattice);
  if (blockIdx.x == 0 && threadIdx.x < 6) ref_lattice[threadIdx.x] = lattice[threadIdx.x];
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 p = posq[i];
    ref[i] = p;
    unsigned key = nb_voxel_key(p, L, voxel_rank, nva, nvb, nvc);
    atom_key[i] = key;
    atomicAdd(&voxel_count[key], 1u);
  }
}

// Exclusive scan of voxel_count -> voxel_off (and cleared scatter
// cursors) over chunks of NB_SCAN_CHUNK voxels: chunk sums, then a
// workgroup scan per chunk.  Up to NB_SCAN_FUSED_MAX chunks (16 M atoms),
// each workgroup sums the chunk sums before its own; beyond that a
// one-workgroup pass turns them into offsets first (k_voxel_scan_parts).
#define NB_SCAN_CHUNK 2048u  // 256 threads x 8 voxels
#define NB_SCAN_FUSED_MAX 4096u

__device__ __forceinline__ unsigned nb_block_excl_scan256(unsigned v, unsigned* tmp,
                                                          unsigned& total) {
  unsigned t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (unsigned off = 1; off < 256u; off <<= 1) {
    unsigned u = (t >= off) ? tmp[t - off] : 0u;
    __syncthreads();
    tmp[t] += u;
    __syncthreads();
  }
  total = tmp[255];
  unsigned r = tmp[t] - v;
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_voxel_scan_reduce(const unsigned* __restrict__ count,
                                                           unsigned* __restrict__ part,
                                                           const unsigned* status, unsigned nvox) {
  if (!nb_rebuild_due(status)) return;
  __shared__ unsigned tmp[256];
  unsigned base = blockIdx.x * NB_SCAN_CHUNK;
  unsigned s = 0;
#pragma unroll
  for (unsigned k = 0; k < NB_SCAN_CHUNK / 256u; ++k) {
    unsigned i = base + k * 256u + threadIdx.x;
    if (i < nvox) s += count[i];
  }
  unsigned total;
  (void)nb_block_excl_scan256(s, tmp, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}


__global__ void __launch_bounds__(1024) k_voxel_scan_parts(unsigned* __restrict__ part,
                                                           const unsigned* status,
                                                           unsigned nparts) {
  if (!nb_rebuild_due(status)) return;
  __shared__ unsigned tmp[1024];
  unsigned t = threadIdx.x, carry = 0;
  for (unsigned base = 0; base < nparts; base += 1024u) {
    unsigned i = base + t;
    unsigned v = (i < nparts) ? part[i] : 0u;
    tmp[t] = v;
    __syncthreads();
    for (unsigned off = 1; off < 1024u; off <<= 1) {
      unsigned u = (t >= off) ? tmp[t - off] : 0u;
      __syncthreads();
      tmp[t] += u;
      __syncthreads();
    }
    if (i < nparts) part[i] = carry + tmp[t] - v;
    carry += tmp[1023];
    __syncthreads();
  }
  if (t == 0) part[nparts] = carry;
}
__global__ void __launch_bounds__(256) k_voxel_scan_apply(const unsigned* __restrict__ count,
                                                          const unsigned* __restrict__ part,
                                                          unsigned* __restrict__ voxel_off,
                                                          unsigned* __restrict__ voxel_cursor,
                                                          const unsigned* status, unsigned nvox,
                                                          unsigned nparts) {
  if (!nb_rebuild_due(status)) return;
  __shared__ unsigned tmp[256];
  const unsigned per = NB_SCAN_CHUNK / 256u;
  unsigned b0 = blockIdx.x * NB_SCAN_CHUNK + threadIdx.x * per;
  unsigned c[NB_SCAN_CHUNK / 256u];
  unsigned s = 0;
#pragma unroll
  for (unsigned k = 0; k < per; ++k) {
    c[k] = (b0 + k < nvox) ? count[b0 + k] : 0u;
    s += c[k];
  }
  unsigned before, all;
  if (nparts <= NB_SCAN_FUSED_MAX) {
    // offset of this chunk: the sum of the chunk sums before it; chunk 0
    // also sums them all, for voxel_off[nvox]
    unsigned lim = blockIdx.x == 0 ? nparts : blockIdx.x;
    unsigned o = 0;
    for (unsigned k = threadIdx.x; k < lim; k += 256u) o += part[k];
    (void)nb_block_excl_scan256(o, tmp, before);
    all = before;
    if (blockIdx.x == 0) before = 0u;
  } else {
    // offsets from k_voxel_scan_parts
    before = part[blockIdx.x];
    all = part[nparts];
  }
  unsigned total;
  unsigned run = before + nb_block_excl_scan256(s, tmp, total);
#pragma unroll
  for (unsigned k = 0; k < per; ++k) {
    if (b0 + k < nvox) {
      voxel_off[b0 + k] = run;
      voxel_cursor[b0 + k] = 0u;
    }
    run += c[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) voxel_off[nvox] = all;
}

__global__ void k_voxel_scatter(const unsigned* __restrict__ atom_key,
                                const unsigned* __restrict__ voxel_off,
                                unsigned* __restrict__ voxel_cursor,
                                unsigned* __restrict__ sorted_id, const unsigned* status,
                                unsigned n) {
  if (!nb_rebuild_due(status)) return;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned k = atom_key[i];
    unsigned slot = voxel_off[k] + atomicAdd(&voxel_cursor[k], 1u);
    sorted_id[slot] = i;
  }
}

__device__ __forceinline__ void nb_sort_voxel(unsigned v, const unsigned* __restrict__ voxel_off,
                                              unsigned* __restrict__ sorted_id,
                                              unsigned* __restrict__ slot_of) {
  unsigned b = voxel_off[v], e = voxel_off[v + 1];
  for (unsigned k = b + 1; k < e; ++k) {
    unsigned key = sorted_id[k];
    int pos = (int)k - 1;
    while (pos >= (int)b && sorted_id[pos] > key) {
      sorted_id[pos + 1] = sorted_id[pos];
      --pos;
    }
    sorted_id[pos + 1] = key;
  }
  for (unsigned k = b; k < e; ++k) slot_of[sorted_id[k]] = k;
}

// Per voxel: sort its atom ids ascending (a voxel holds a handful of
// atoms) so the slot order is independent of atomic arrival order; then
// clear the voxel histogram for the next rebuild and record slot_of[id].
__global__ void k_voxel_sort(const unsigned* __restrict__ voxel_off,
                             unsigned* __restrict__ voxel_count, unsigned* __restrict__ sorted_id,
                             unsigned* __restrict__ slot_of, const unsigned* status,
                             unsigned nvox) {
  if (!nb_rebuild_due(status)) return;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += gridDim.x * blockDim.x) {
    nb_sort_voxel(v, voxel_off, sorted_id, slot_of);
    voxel_count[v] = 0u;
  }
}

// Per-slot static data refreshed at a rebuild: type and exclusion key in
// slot order, padding lanes marked with sentinels.

// Every step: gather positions into slot order (wrapped into the primary
// cell) and clear the j-side fixed-point accumulators of the pair kernel.
// Slot-order copy of the positions (wrapped into the cell) and clearing of
// the j-side accumulators, every step.  On rebuild steps (status given and
// the request set) it also fills the slots' types and keys and each 64-slot
// block's bounding box (one wave per block), from the same positions.
// One slot s < npad; the wave's 64 lanes hold the 64 slots of one block
// (the bounding-box reduction is a wave reduction).
__device__ __forceinline__ void nb_gather_slot(unsigned s, bool rebuild,
                                               const float4* __restrict__ posq,
                                               const unsigned* __restrict__ sorted_id,
                                               const float* __restrict__ lattice,
                                               float4* __restrict__ tile_posq,
                                               long long* __restrict__ fpj,
                                               long long* __restrict__ fpj_ew, unsigned n,
                                               const unsigned* __restrict__ type,
                                               const unsigned* __restrict__ key,
                                               unsigned* __restrict__ tile_type,
                                               unsigned* __restrict__ tile_key,
                                               float4* __restrict__ center,
                                               float4* __restrict__ half) {
  float4 out;
  unsigned id = 0u;
  if (s < n) {
    HdLattice L = hd_load_lattice(lattice);
    id = sorted_id[s];
    float4 p = posq[id];
    int a, b, c;
    hd_wrap_with_image(p.x, p.y, p.z, a, b, c, L);
    out = p;
  } else {
    out = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  tile_posq[s] = out;
  longlong2 z = make_longlong2(0, 0);
  reinterpret_cast<longlong2*>(fpj)[2 * s] = z;
  reinterpret_cast<longlong2*>(fpj)[2 * s + 1] = z;
  if (fpj_ew) reinterpret_cast<longlong2*>(fpj_ew)[s] = z;
  if (!rebuild) return;
  bool v = s < n;
  tile_type[s] = v ? type[id] : 0u;
  tile_key[s] = v ? key[id] : HD_SENTINEL - 1u;
  float lox = v ? out.x : 3.0e38f, loy = v ? out.y : 3.0e38f, loz = v ? out.z : 3.0e38f;
  float hix = v ? out.x : -3.0e38f, hiy = v ? out.y : -3.0e38f, hiz = v ? out.z : -3.0e38f;
  for (int off = 32; off > 0; off >>= 1) {
    lox = fminf(lox, __shfl_xor(lox, off));
    loy = fminf(loy, __shfl_xor(loy, off));
    loz = fminf(loz, __shfl_xor(loz, off));
    hix = fmaxf(hix, __shfl_xor(hix, off));
    hiy = fmaxf(hiy, __shfl_xor(hiy, off));
    hiz = fmaxf(hiz, __shfl_xor(hiz, off));
  }
  if ((threadIdx.x & 63u) == 0) {
    unsigned blk = s >> 6;
    center[blk] = make_float4(0.5f * (lox + hix), 0.5f * (loy + hiy), 0.5f * (loz + hiz), 0.f);
    half[blk] = make_float4(0.5f * (hix - lox), 0.5f * (hiy - loy), 0.5f * (hiz - loz), 0.f);
  }
}

__global__ void k_tile_gather(const float4* __restrict__ posq, const unsigned* __restrict__ sorted_id,
                              const float* __restrict__ lattice, float4* __restrict__ tile_posq,
                              long long* __restrict__ fpj, long long* __restrict__ fpj_ew,
                              unsigned n, unsigned npad, const unsigned* status,
                              const unsigned* __restrict__ type, const unsigned* __restrict__ key,
                              unsigned* __restrict__ tile_type, unsigned* __restrict__ tile_key,
                              float4* __restrict__ center, float4* __restrict__ half) {
  unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npad) return;  // whole waves: npad is a multiple of 64
  nb_gather_slot(s, status && nb_rebuild_due(status), posq, sorted_id, lattice, tile_posq, fpj,
                 fpj_ew, n, type, key, tile_type, tile_key, center, half);
}


// Per-block axis-aligned bounding box: centre and half extents (one wave
// per block).  Padding lanes are ignored.

// Closest approach of two boxes under the minimum image (per axis; exact
// for orthorhombic cells), as kernels/neighbor.cu:364-388.  In a tilted
// cell a pair's own minimum image (fractional rounding of its separation)
// can differ by one lattice vector per axis from the one the box centres
// round to, so the 27 images around the centres' are all tried there.
__device__ __forceinline__ float nb_box_gap2(float dx, float dy, float dz, float4 ha, float4 hb) {
  float ex = fmaxf(fabsf(dx) - (ha.x + hb.x), 0.f);
  float ey = fmaxf(fabsf(dy) - (ha.y + hb.y), 0.f);
  float ez = fmaxf(fabsf(dz) - (ha.z + hb.z), 0.f);
  return ex * ex + ey * ey + ez * ez;
}
__device__ __forceinline__ float nb_box_dist2(float4 ca, float4 ha, float4 cb, float4 hb,
                                              const HdLattice& L) {
  float dx = ca.x - cb.x, dy = ca.y - cb.y, dz = ca.z - cb.z;
  hd_min_image(dx, dy, dz, L);
  if (L.xy == 0.f && L.xz == 0.f && L.yz == 0.f) return nb_box_gap2(dx, dy, dz, ha, hb);
  float best = 3.0e38f;
  for (int kc = -1; kc <= 1; ++kc)
    for (int kb = -1; kb <= 1; ++kb)
      for (int ka = -1; ka <= 1; ++ka) {
        float sx = dx + (float)ka * L.lx + (float)kb * L.xy + (float)kc * L.xz;
        float sy = dy + (float)kb * L.ly + (float)kc * L.yz;
        float sz = dz + (float)kc * L.lz;
        best = fminf(best, nb_box_gap2(sx, sy, sz, ha, hb));
      }
  return best;
}

// Block-pair candidates: bit (jb % 64) of mask[b * nchunks + jb / 64] is
// set when j-block jb's box lies within r_search of block b's (jb >= b only
// for the half list).  Each wave handles (b, chunk of 64 j-blocks) items in
// a grid-stride loop, so the box tests of a rebuild run at full width
// instead of as a dependent loop inside each block's wave, while the launch
// stays small on the steps where the rebuild gate closes it at once.
#define NB_PAIR_WAVES 4096
// One (block, chunk of 64 j-blocks) item, one wave.
__device__ __forceinline__ void nb_block_pair_item(unsigned wv, const float4* __restrict__ center,
                                                   const float4* __restrict__ half,
                                                   const HdLattice& L, float rs2,
                                                   unsigned long long* __restrict__ mask,
                                                   unsigned nblocks, unsigned nchunks, int full) {
  unsigned lane = threadIdx.x & 63;
  unsigned b = wv / nchunks, c = wv % nchunks;
  unsigned jb = c * 64 + lane;
  bool cand = false;
  if (jb < nblocks && (full || jb >= b))
    cand = nb_box_dist2(center[b], half[b], center[jb], half[jb], L) <= rs2;
  unsigned long long m = __ballot(cand);
  if (lane == 0) mask[wv] = m;
}
__global__ void __launch_bounds__(256) k_block_pairs(const float4* __restrict__ center,
                                                     const float4* __restrict__ half,
                                                     const float* __restrict__ lattice, float rs2,
                                                     unsigned long long* __restrict__ mask,
                                                     unsigned nblocks, unsigned nchunks, int full,
                                                     const unsigned* status) {
  if (!nb_rebuild_due(status)) return;
  unsigned nwaves = gridDim.x * (blockDim.x >> 6);
  HdLattice L = hd_load_lattice(lattice);
  for (unsigned wv = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; wv < nblocks * nchunks;
       wv += nwaves)
    nb_block_pair_item(wv, center, half, L, rs2, mask, nblocks, nchunks, full);
}

// Systems of at most NB_SMALL_ATOMS atoms and NB_SMALL_VOX voxels: the
// displacement check, the rebuild's voxel binning (keys, histogram, scan,
// scatter, per-voxel sort), the tile gather (with the SPME spline records
// when asked) and the block-pair box tests run in one workgroup, with
// barriers between the phases in place of the nine launches of the general
// path.  The histogram and the scatter cursors live in LDS (the
// global voxel_count stays zero, as the general path leaves it).  Results
// are those of the general path: the per-atom arithmetic is shared with its
// kernels, the scan is exact and the per-voxel sort makes the slot order
// independent of the order of the atomics.
#define NB_SMALL_THREADS 1024u
#define NB_SMALL_ATOMS 8192u
#define NB_SMALL_VOX 8192u

__global__ void __launch_bounds__(NB_SMALL_THREADS) k_nb_small(HdNeighborArgs a, int check,
                                                               long long* __restrict__ fpj,
                                                               long long* __restrict__ fpj_ew,
                                                               float4* __restrict__ spl,
                                                               unsigned* __restrict__ sg,
                                                               unsigned gna, unsigned gnb,
                                                               unsigned gnc) {
  __shared__ unsigned cnt[NB_SMALL_VOX];
  __shared__ unsigned scan[NB_SMALL_THREADS];
  __shared__ int s_over;
  const unsigned t = threadIdx.x;
  HdLattice L = hd_load_lattice(a.lattice);
  if (t == 0) s_over = 0;
  __syncthreads();
  if (check) {
    // k_disp_check's test per atom
    float dl = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) dl = fmaxf(dl, fabsf(a.lattice[k] - a.ref_lattice[k]));
    bool over = false;
    for (unsigned i = t; i < a.n; i += NB_SMALL_THREADS) {
      float4 p = a.posq[i], r = a.ref_pos[i];
      float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
      hd_min_image(dx, dy, dz, L);
      over = over || sqrtf(dx * dx + dy * dy + dz * dz) + 1.7320508f * dl > a.half_skin;
    }
    if (__any(over) && (t & 63u) == 0) {
      atomicOr(a.status, NB_STATUS_REBUILD);
      s_over = 1;
    }
  }
  __syncthreads();
  // one thread decides and every thread reads its decision, so the
  // barriers of the rebuild phases are reached by the whole workgroup
  if (t == 0) s_over = (s_over != 0 || nb_rebuild_due(a.status)) ? 1 : 0;
  __syncthreads();
  const bool rebuild = s_over != 0;
  if (rebuild) {
    if (t < 6) a.ref_lattice[t] = a.lattice[t];
    for (unsigned v = t; v < a.nvox; v += NB_SMALL_THREADS) cnt[v] = 0u;
    __syncthreads();
    for (unsigned i = t; i < a.n; i += NB_SMALL_THREADS) {
      float4 p = a.posq[i];
      a.ref_pos[i] = p;
      unsigned key = nb_voxel_key(p, L, a.voxel_rank, a.nv[0], a.nv[1], a.nv[2]);
      a.atom_key[i] = key;
      atomicAdd(&cnt[key], 1u);
    }
    __syncthreads();
    // exclusive scan: thread t owns voxels [t per, (t + 1) per)
    const unsigned per = (a.nvox + NB_SMALL_THREADS - 1u) / NB_SMALL_THREADS;
    const unsigned b0 = t * per;
    unsigned sum = 0u;
    for (unsigned k = 0; k < per; ++k)
      if (b0 + k < a.nvox) sum += cnt[b0 + k];
    scan[t] = sum;
    __syncthreads();
    for (unsigned off = 1; off < NB_SMALL_THREADS; off <<= 1) {
      unsigned u = t >= off ? scan[t - off] : 0u;
      __syncthreads();
      scan[t] += u;
      __syncthreads();
    }
    unsigned run = scan[t] - sum;
    for (unsigned k = 0; k < per; ++k) {
      unsigned v = b0 + k;
      if (v < a.nvox) {
        unsigned c = cnt[v];
        a.voxel_off[v] = run;
        cnt[v] = run;  // the scatter cursor
        run += c;
      }
    }
    if (t == NB_SMALL_THREADS - 1u) a.voxel_off[a.nvox] = scan[t];
    __syncthreads();
    for (unsigned i = t; i < a.n; i += NB_SMALL_THREADS)
      a.sorted_id[atomicAdd(&cnt[a.atom_key[i]], 1u)] = i;
    __syncthreads();
    for (unsigned v = t; v < a.nvox; v += NB_SMALL_THREADS)
      nb_sort_voxel(v, a.voxel_off, a.sorted_id, a.slot_of);
    __syncthreads();
  }
  // npad is a multiple of 64 and so is the stride: each wave covers whole
  // blocks in every pass
  for (unsigned s = t; s < a.npad; s += NB_SMALL_THREADS)
    nb_gather_slot(s, rebuild, a.posq, a.sorted_id, a.lattice, a.tile_posq, fpj, fpj_ew, a.n,
                   a.type, a.key, a.tile_type, a.tile_key, a.center, a.half);
  if (spl) {
    // SPME spline records from the gathered positions (each thread reads
    // back the slots it wrote)
    for (unsigned s = t; s < a.n; s += NB_SMALL_THREADS)
      hd_spline_record(a.tile_posq[s], L, gna, gnb, gnc, spl, sg, s);
  }
  if (!rebuild) return;
  __syncthreads();
  // the box tests read every block's bounding box from the gather
  const unsigned nblocks = a.npad / 64u, nchunks = (nblocks + 63u) / 64u;
  for (unsigned wv = t >> 6; wv < nblocks * nchunks; wv += NB_SMALL_THREADS / 64u)
    nb_block_pair_item(wv, a.center, a.half, L, a.rs2, a.block_mask, nblocks, nchunks,
                       a.atom_lists);
}

bool hd_neighbor_small(const HdNeighborArgs& a) {
  return a.n <= NB_SMALL_ATOMS && a.nvox <= NB_SMALL_VOX;
}

void hd_launch_neighbor_small(hipStream_t st, const HdNeighborArgs& a, bool check, long long* fpj,
                              long long* fpj_ew, float4* spl, unsigned* sg, unsigned na,
                              unsigned nb, unsigned nc) {
  hipLaunchKernelGGL(k_nb_small, dim3(1), dim3(NB_SMALL_THREADS), 0, st, a, check ? 1 : 0, fpj,
                     fpj_ew, spl, sg, na, nb, nc);
}

#define NB_CAND_MAX 1024

// The candidate j-blocks of block b in increasing order, from the chunk
// masks, into an LDS list (at most NB_CAND_MAX; *too_many flags more).
__device__ __forceinline__ unsigned nb_collect(const unsigned long long* __restrict__ mask,
                                               unsigned b, unsigned nchunks, unsigned lane,
                                               unsigned* cand, bool* too_many) {
  unsigned ncand = 0;
  for (unsigned c0 = 0; c0 < nchunks; c0 += 64) {
    unsigned c = c0 + lane;
    unsigned long long mk = c < nchunks ? mask[(size_t)b * nchunks + c] : 0ull;
    unsigned cnt = __popcll(mk);
    unsigned inc = cnt;
#pragma unroll
    for (unsigned off = 1; off < 64; off <<= 1) {
      unsigned v = __shfl_up(inc, off);
      if (lane >= off) inc += v;
    }
    unsigned base = ncand + inc - cnt;
    while (mk) {
      unsigned bit = __builtin_ctzll(mk);
      mk &= mk - 1;
      if (base < NB_CAND_MAX) cand[base] = c * 64 + bit;
      ++base;
    }
    ncand += __shfl(inc, 63);
  }
  __builtin_amdgcn_wave_barrier();
  if (ncand > NB_CAND_MAX) {
    *too_many = true;
    ncand = NB_CAND_MAX;
  }
  return ncand;
}

// Build the packed list: one wave per i-block.  The candidate j-blocks
// (>= b, increasing order) are streamed through the wave one block per step,
// the next block's atoms loaded while the current one is handled.  A j-atom
// farther than r_search from the i-block's box cannot hit any i-atom; the
// others are queued in LDS, and every time 64 are queued each lane tests
// one queued j-atom against the 64 i-atoms.  Queued j-atoms with at least
// one hit are appended, in queue order, to a staging buffer flushed 64 at a
// time.  The entries are therefore the hit j-atoms in (j-block, slot) order:
// a pure function of the positions.
#define NB_WAVES 4
#define NB_CHUNK 4  // candidate blocks whose loads are in flight together

// A j-atom's offset from the i-block centre in the image nearest it, as
// hd_rel_to_center computes it (FP64 difference less whole box lengths,
// rounded once), with the image count from an f32 product with the
// reciprocal box length.  The counts can differ only for an offset within
// an ulp of half a box length, and such an atom is rejected by the box
// test in either image (the block lies at least r_search inside the half
// box on this path).
__device__ __forceinline__ void nb_rel_to_center(float x, float y, float z, float4 cb,
                                                 const HdLattice& L, float ilx, float ily,
                                                 float ilz, float& rx, float& ry, float& rz) {
  double nx = (double)rintf((x - cb.x) * ilx);
  double ny = (double)rintf((y - cb.y) * ily);
  double nz = (double)rintf((z - cb.z) * ilz);
  rx = (float)(((double)x - (double)cb.x) - (double)L.lx * nx);
  ry = (float)(((double)y - (double)cb.y) - (double)L.ly * ny);
  rz = (float)(((double)z - (double)cb.z) - (double)L.lz * nz);
}

// one pass of the 64 x 64 test over the queued j-atoms (lane = queue
// position); returns whether this lane's j-atom hits any i-atom
__device__ __forceinline__ bool nb_test_queued(const float4* ipos, float4 pj, unsigned jslot,
                                               bool jvalid, bool rel, unsigned b, unsigned ni,
                                               float rs2, const HdLattice& L) {
  bool hit = false;
  if (!jvalid) return false;
  if (rel) {
#pragma unroll 8
    for (unsigned m = 0; m < 64; ++m) {
      float4 a = ipos[m];
      float dx = a.x - pj.x, dy = a.y - pj.y, dz = a.z - pj.z;
      float r2 = dx * dx + dy * dy + dz * dz;
      hit |= (r2 <= rs2) & (m < ni) & ((b * 64 + m) != jslot);
    }
  } else {
#pragma unroll 8
    for (unsigned m = 0; m < 64; ++m) {
      float4 a = ipos[m];
      float dx = a.x - pj.x, dy = a.y - pj.y, dz = a.z - pj.z;
      hd_min_image(dx, dy, dz, L);
      float r2 = dx * dx + dy * dy + dz * dz;
      hit |= (r2 <= rs2) & (m < ni) & ((b * 64 + m) != jslot);
    }
  }
  return hit;
}

#define NB_FOLD_MAX_BLOCKS 128u  // block counts the folded ordering handles

// FOLD (at most NB_FOLD_MAX_BLOCKS blocks): the workgroup that takes the
// last ticket (*ticket, 0 between launches) also does k_rebuild_clear's
// work once every block's entry count is written: blocks ordered by entry
// count, heaviest first (ties by block index), the rebuild request cleared
// and the rebuild counted.  The order only maps workgroups to blocks in
// the pair kernel, so any permutation gives the same forces.
template <bool FOLD>
__global__ void __launch_bounds__(256) k_build_entries(
    const float4* __restrict__ tile_posq, const float4* __restrict__ center,
    const float4* __restrict__ half, const float* __restrict__ lattice, float rs2,
    const unsigned long long* __restrict__ mask, unsigned nchunks,
    unsigned* __restrict__ entry_atoms, unsigned* __restrict__ entry_count, unsigned cap,
    unsigned* __restrict__ max_count, unsigned* status, unsigned n, unsigned nblocks,
    unsigned* ticket, unsigned* rebuild_count, unsigned* __restrict__ order) {
  if (!nb_rebuild_due(status)) return;
  __shared__ float4 ipos[NB_WAVES][64];
  __shared__ unsigned buf[NB_WAVES][128];
  __shared__ float4 qpos[NB_WAVES][128];
  __shared__ unsigned qslot[NB_WAVES][128];
  __shared__ unsigned cand[NB_WAVES][NB_CAND_MAX];
  unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned b = blockIdx.x * NB_WAVES + w;
  if (b < nblocks) {
  HdLattice L = hd_load_lattice(lattice);
  unsigned islot = b * 64 + lane;
  float4 pi = tile_posq[islot];
  if (islot >= n) pi = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.f);
  ipos[w][lane] = pi;
  __builtin_amdgcn_wave_barrier();
  float4 cb = center[b], hb = half[b];
  unsigned ni = min(64u, n - b * 64);
  // When the block is far from every periodic boundary (orthorhombic box,
  // L/2 - half extent >= r_search), hold i and j relative to the block
  // centre with the periodic shift applied once per atom, so each of the
  // 64 x 64 distance tests is plain arithmetic (no min-image, no division).
  float rsearch = sqrtf(rs2);
  bool rel = (L.xy == 0.f && L.xz == 0.f && L.yz == 0.f) && (0.5f * L.lx - hb.x >= rsearch) &&
             (0.5f * L.ly - hb.y >= rsearch) && (0.5f * L.lz - hb.z >= rsearch);
  if (rel) {
    float4 r;
    hd_rel_to_center(pi.x, pi.y, pi.z, cb.x, cb.y, cb.z, L, r.x, r.y, r.z);
    r.w = 0.f;
    ipos[w][lane] = r;
  }
  bool too_many = false;
  unsigned ncand = nb_collect(mask, b, nchunks, lane, cand[w], &too_many);
  __builtin_amdgcn_wave_barrier();
  const float filx = 1.0f / L.lx, fily = 1.0f / L.ly, filz = 1.0f / L.lz;
  unsigned nbuf = 0, nent = 0, nq = 0;
  unsigned* out = entry_atoms + (size_t)b * cap * 64;
  // test 64 queued atoms at a time (the remainder after the last block)
  auto drain = [&](bool last) {
    while (nq >= 64 || (last && nq > 0)) {
      bool qvalid = lane < nq;
      float4 qp = qpos[w][lane];
      unsigned qs = qvalid ? qslot[w][lane] : HD_SENTINEL;
      bool hit = nb_test_queued(ipos[w], qp, qs, qvalid, rel, b, ni, rs2, L);
      unsigned rem = nq > 64 ? nq - 64 : 0;
      float4 rp = qpos[w][64 + (lane < rem ? lane : 0)];
      unsigned rsl = qslot[w][64 + (lane < rem ? lane : 0)];
      __builtin_amdgcn_wave_barrier();
      if (lane < rem) {
        qpos[w][lane] = rp;
        qslot[w][lane] = rsl;
      }
      nq = rem;
      unsigned long long hmask = __ballot(hit);
      if (hit) {
        unsigned pos = nbuf + __popcll(hmask & hd_lane_mask_lt());
        buf[w][pos] = qs;
      }
      nbuf += __popcll(hmask);
      __builtin_amdgcn_wave_barrier();
      if (nbuf >= 64) {
        if (nent < cap) out[(size_t)nent * 64 + lane] = buf[w][lane];
        ++nent;
        unsigned brem = nbuf - 64;
        unsigned v = (lane < brem) ? buf[w][64 + lane] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (lane < brem) buf[w][lane] = v;
        nbuf = brem;
        __builtin_amdgcn_wave_barrier();
      }
    }
  };
  // the candidate blocks' atoms are loaded NB_CHUNK blocks at a time, the
  // next group while the current one is queued and tested
  float4 pn[NB_CHUNK];
#pragma unroll
  for (unsigned u = 0; u < NB_CHUNK; ++u)
    pn[u] = (u < ncand) ? tile_posq[cand[w][u] * 64 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
  for (unsigned c0 = 0; c0 < ncand; c0 += NB_CHUNK) {
    float4 pc[NB_CHUNK];
#pragma unroll
    for (unsigned u = 0; u < NB_CHUNK; ++u) {
      pc[u] = pn[u];
      unsigned cn = c0 + NB_CHUNK + u;
      if (cn < ncand) pn[u] = tile_posq[cand[w][cn] * 64 + lane];
    }
#pragma unroll
    for (unsigned u = 0; u < NB_CHUNK; ++u) {
      if (c0 + u >= ncand) break;
      unsigned jb = cand[w][c0 + u];
      unsigned jslot = jb * 64 + lane;
      bool jvalid = jslot < n;
      float4 pj = pc[u];
      // distance from the j-atom to the i-block's box
      float dx, dy, dz;
      if (rel) {
        nb_rel_to_center(pj.x, pj.y, pj.z, cb, L, filx, fily, filz, dx, dy, dz);
        pj.x = dx; pj.y = dy; pj.z = dz;
      } else {
        dx = pj.x - cb.x; dy = pj.y - cb.y; dz = pj.z - cb.z;
        hd_min_image(dx, dy, dz, L);
      }
      float ex = fmaxf(fabsf(dx) - hb.x, 0.f), ey = fmaxf(fabsf(dy) - hb.y, 0.f),
            ez = fmaxf(fabsf(dz) - hb.z, 0.f);
      bool near = jvalid && (ex * ex + ey * ey + ez * ez <= rs2);
      unsigned long long nmask = __ballot(near);
      if (near) {
        unsigned pos = nq + __popcll(nmask & hd_lane_mask_lt());
        qpos[w][pos] = pj;
        qslot[w][pos] = jslot;
      }
      nq += __popcll(nmask);
      __builtin_amdgcn_wave_barrier();
      drain(false);
    }
  }
  drain(true);
  if (nbuf > 0) {
    if (nent < cap) out[(size_t)nent * 64 + lane] = (lane < nbuf) ? buf[w][lane] : HD_SENTINEL;
    ++nent;
  }
  if (lane == 0) {
    entry_count[b] = min(nent, cap);
    atomicMax(max_count, nent);
    if (nent > cap || too_many) atomicOr(status, NB_STATUS_OVERFLOW);
    else if (4 * nent > 3 * cap) atomicOr(status, NB_STATUS_HIGHWATER);
  }
  }  // b < nblocks
  if (!FOLD) return;
  __shared__ unsigned s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(ticket, 1u) == gridDim.x - 1u ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // every block's counts and status bits are visible
  __shared__ unsigned s_cnt[NB_FOLD_MAX_BLOCKS];
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x) s_cnt[i] = entry_count[i];
  __syncthreads();
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x) {
    unsigned ci = s_cnt[i], pos = 0;
    for (unsigned j = 0; j < nblocks; ++j) {
      unsigned cj = s_cnt[j];
      pos += (cj > ci || (cj == ci && j < i)) ? 1u : 0u;
    }
    order[pos] = i;
  }
  if (threadIdx.x == 0) {
    atomicAnd(status, ~(NB_STATUS_REBUILD | NB_STATUS_FORCE));
    atomicAdd(rebuild_count, 1u);
    atomicExch(ticket, 0u);
  }
}

// Per-atom full lists (the atom-list mode): one wave per i-block, lane =
// i-atom.  The wave reads its candidate j-blocks (every j-block whose box
// lies within r_search, not only j >= i) into an LDS list, then stages the
// candidates' atoms in LDS one block at a time, loading the next block
// while it tests the current one.  Each lane appends the j-atoms within
// r_search of its own atom that are neither itself nor excluded by the key
// rule.  Row k of block b holds the k-th neighbour of every lane
// (column-major, so the pair kernel's index loads are coalesced); rows past
// a lane's count up to the block's longest list hold HD_SENTINEL.  Each
// pair appears once in each atom's list, so the pair kernel needs no j-side
// accumulation.  The order is j-block by j-block, then j-slot: the lists
// are a pure function of the positions.

__device__ __forceinline__ void alist_stage(const float4* __restrict__ tile_posq,
                                            const unsigned* __restrict__ tile_key, unsigned jb,
                                            unsigned lane, unsigned n, float4& p, unsigned& key) {
  unsigned jslot = jb * 64 + lane;
  bool jvalid = jslot < n;
  p = tile_posq[jvalid ? jslot : 0u];
  key = jvalid ? tile_key[jslot] : 0u;
}

__global__ void __launch_bounds__(256) k_build_alist(
    const float4* __restrict__ tile_posq, const unsigned* __restrict__ tile_key,
    unsigned excl_span, const float4* __restrict__ center, const float4* __restrict__ half,
    const float* __restrict__ lattice, float rs2, const unsigned long long* __restrict__ mask,
    unsigned nchunks, unsigned* __restrict__ rows_out, unsigned* __restrict__ row_count,
    unsigned cap, unsigned* __restrict__ max_count, unsigned* status, unsigned n,
    unsigned nblocks) {
  if (!nb_rebuild_due(status)) return;
  __shared__ float4 jpos[NB_WAVES][64];
  __shared__ unsigned jkey[NB_WAVES][64];
  __shared__ unsigned cand[NB_WAVES][NB_CAND_MAX];
  unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned b = blockIdx.x * NB_WAVES + w;
  if (b >= nblocks) return;
  HdLattice L = hd_load_lattice(lattice);
  unsigned islot = b * 64 + lane;
  bool ivalid = islot < n;
  float4 pi = tile_posq[ivalid ? islot : 0u];
  unsigned ikey = ivalid ? tile_key[islot] : 0u;
  float4 cb = center[b], hb = half[b];
  float rsearch = sqrtf(rs2);
  // far from every periodic boundary (orthorhombic box): i and j relative
  // to the block centre, the periodic shift applied once per staged atom
  bool rel = (L.xy == 0.f && L.xz == 0.f && L.yz == 0.f) && (0.5f * L.lx - hb.x >= rsearch) &&
             (0.5f * L.ly - hb.y >= rsearch) && (0.5f * L.lz - hb.z >= rsearch);
  float ix = pi.x, iy = pi.y, iz = pi.z;
  if (rel) hd_rel_to_center(pi.x, pi.y, pi.z, cb.x, cb.y, cb.z, L, ix, iy, iz);
  bool too_many = false;
  unsigned ncand = nb_collect(mask, b, nchunks, lane, cand[w], &too_many);
  const double ilx = 1.0 / (double)L.lx, ily = 1.0 / (double)L.ly, ilz = 1.0 / (double)L.lz;
  unsigned cnt = 0;
  unsigned* out = rows_out + (size_t)b * cap * 64;
  float4 pn = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned kn = 0;
  if (ncand) alist_stage(tile_posq, tile_key, cand[w][0], lane, n, pn, kn);
  for (unsigned c = 0; c < ncand; ++c) {
    unsigned jb = cand[w][c];
    float4 pj = pn;
    if (rel) {
      float jx, jy, jz;
      hd_rel_to_center_rcp(pj.x, pj.y, pj.z, cb.x, cb.y, cb.z, L, ilx, ily, ilz, jx, jy, jz);
      pj.x = jx; pj.y = jy; pj.z = jz;
    }
    jpos[w][lane] = pj;
    jkey[w][lane] = kn;
    __builtin_amdgcn_wave_barrier();
    if (c + 1 < ncand) alist_stage(tile_posq, tile_key, cand[w][c + 1], lane, n, pn, kn);
    unsigned nj = min(64u, n - jb * 64);
    if (ivalid) {
#pragma unroll 8
      for (unsigned m = 0; m < 64; ++m) {
        float4 a = jpos[w][m];
        float dx = ix - a.x, dy = iy - a.y, dz = iz - a.z;
        if (!rel) hd_min_image(dx, dy, dz, L);
        float r2 = dx * dx + dy * dy + dz * dz;
        unsigned js = jb * 64 + m;
        bool ok = (jkey[w][m] - ikey + excl_span) > 2u * excl_span;
        if (r2 <= rs2 && ok && js != islot && m < nj) {
          if (cnt < cap) out[(size_t)cnt * 64 + lane] = js;
          ++cnt;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  unsigned mx = cnt;
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, off));
  unsigned rows = min(mx, cap);
  for (unsigned k = cnt; k < rows; ++k) out[(size_t)k * 64 + lane] = HD_SENTINEL;
  if (lane == 0) {
    row_count[b] = rows;
    atomicMax(max_count, mx);
    if (mx > cap || too_many) atomicOr(status, NB_STATUS_OVERFLOW);
    else if (4 * mx > 3 * cap) atomicOr(status, NB_STATUS_HIGHWATER);
  }
}

// End of a rebuild: remember the reference positions and clear the
// rebuild request bits (sticky overflow/high-water bits stay for the host).

// Clears the rebuild request and, on rebuild steps, orders the blocks by
// entry count, heaviest first, as the pair kernel's workgroup order: block
// costs vary about twofold, and a heavy block dispatched last runs alone at
// the end of the launch.  One workgroup: a histogram of the counts, suffix
// sums, a scatter (the order among equal counts is arbitrary; the forces do
// not depend on it, their sums being integer).
#define NB_ORDER_BINS 1024u
__global__ void __launch_bounds__(NB_ORDER_BINS) k_rebuild_clear(
    unsigned* status, unsigned* rebuild_count, const unsigned* __restrict__ entry_count,
    unsigned nblocks, unsigned* __restrict__ order) {
  __shared__ unsigned hist[NB_ORDER_BINS], cur[NB_ORDER_BINS];
  unsigned s = *status;  // read by every thread before thread 0 rewrites it
  if (!(s & (NB_STATUS_REBUILD | NB_STATUS_FORCE))) return;
  unsigned t = threadIdx.x;
  hist[t] = 0u;
  __syncthreads();
  for (unsigned i = t; i < nblocks; i += NB_ORDER_BINS)
    atomicAdd(&hist[min(entry_count[i], NB_ORDER_BINS - 1u)], 1u);
  __syncthreads();
  // hist[c] <- number of blocks whose count is >= c
  for (unsigned d = 1; d < NB_ORDER_BINS; d <<= 1) {
    unsigned v = (t + d < NB_ORDER_BINS) ? hist[t + d] : 0u;
    __syncthreads();
    hist[t] += v;
    __syncthreads();
  }
  // the blocks with count c take positions [hist[c + 1], hist[c])
  cur[t] = (t + 1 < NB_ORDER_BINS) ? hist[t + 1] : 0u;
  __syncthreads();
  for (unsigned i = t; i < nblocks; i += NB_ORDER_BINS) {
    unsigned pos = atomicAdd(&cur[min(entry_count[i], NB_ORDER_BINS - 1u)], 1u);
    order[pos] = i;
  }
  if (t == 0) {
    *status = s & ~(NB_STATUS_REBUILD | NB_STATUS_FORCE);
    *rebuild_count += 1u;
  }
}

// ---------------------------------------------------------------------
// Host launch wrappers
// ---------------------------------------------------------------------
static inline unsigned cdiv(unsigned a, unsigned b) { return (a + b - 1) / b; }

unsigned hd_voxel_scan_parts(unsigned nvox) { return cdiv(nvox, NB_SCAN_CHUNK); }

// Kernels gated by the rebuild flag run grid-stride loops over a modest grid
// (NB_GATED_WGS workgroups): on the steps where the gate closes them they
// retire quickly, and on rebuild steps the loops cover the work.
#define NB_GATED_WGS 512u

void hd_launch_disp_check(hipStream_t st, const HdNeighborArgs& a) {
  const unsigned T = 256;
  hipLaunchKernelGGL(k_disp_check, dim3(cdiv(a.n, T) ? cdiv(a.n, T) : 1), dim3(T), 0, st,
                     a.posq, a.ref_pos, a.lattice, a.ref_lattice, a.half_skin, a.status, a.n);
}

void hd_launch_neighbor_update(hipStream_t st, const HdNeighborArgs& a, bool check) {
  const unsigned T = 256;
  unsigned gridN = min(cdiv(a.n, T), NB_GATED_WGS);
  if (gridN == 0) gridN = 1;
  if (check)
    hipLaunchKernelGGL(k_disp_check, dim3(cdiv(a.n, T) ? cdiv(a.n, T) : 1), dim3(T), 0, st,
                       a.posq, a.ref_pos, a.lattice, a.ref_lattice, a.half_skin, a.status, a.n);
  hipLaunchKernelGGL(k_voxel_keys, dim3(gridN), dim3(T), 0, st, a.posq, a.lattice, a.voxel_rank,
                     a.nv[0], a.nv[1], a.nv[2], a.atom_key, a.voxel_count, a.ref_pos,
                     a.ref_lattice, a.status, a.n);
  unsigned nparts = hd_voxel_scan_parts(a.nvox);
  hipLaunchKernelGGL(k_voxel_scan_reduce, dim3(nparts), dim3(256), 0, st, a.voxel_count,
                     a.voxel_part, a.status, a.nvox);
  if (nparts > NB_SCAN_FUSED_MAX)
    hipLaunchKernelGGL(k_voxel_scan_parts, dim3(1), dim3(1024), 0, st, a.voxel_part, a.status,
                       nparts);
  hipLaunchKernelGGL(k_voxel_scan_apply, dim3(nparts), dim3(256), 0, st, a.voxel_count,
                     a.voxel_part, a.voxel_off, a.voxel_cursor, a.status, a.nvox, nparts);
  hipLaunchKernelGGL(k_voxel_scatter, dim3(gridN), dim3(T), 0, st, a.atom_key, a.voxel_off,
                     a.voxel_cursor, a.sorted_id, a.status, a.n);
  unsigned gridV = min(cdiv(a.nvox, T), NB_GATED_WGS);
  hipLaunchKernelGGL(k_voxel_sort, dim3(gridV), dim3(T), 0, st, a.voxel_off, a.voxel_count,
                     a.sorted_id, a.slot_of, a.status, a.nvox);
}

void hd_launch_tile_gather(hipStream_t st, const float4* posq, const unsigned* sorted_id,
                           const float* lattice, float4* tile_posq, long long* fpj,
                           long long* fpj_ew, unsigned n, unsigned npad) {
  hipLaunchKernelGGL(k_tile_gather, dim3(cdiv(npad, 256)), dim3(256), 0, st, posq, sorted_id,
                     lattice, tile_posq, fpj, fpj_ew, n, npad, nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr);
}

void hd_launch_tile_gather_rebuild(hipStream_t st, const HdNeighborArgs& a, long long* fpj,
                                   long long* fpj_ew) {
  hipLaunchKernelGGL(k_tile_gather, dim3(cdiv(a.npad, 256)), dim3(256), 0, st, a.posq,
                     a.sorted_id, a.lattice, a.tile_posq, fpj, fpj_ew, a.n, a.npad, a.status,
                     a.type, a.key, a.tile_type, a.tile_key, a.center, a.half);
}

unsigned hd_block_pair_chunks(unsigned nblocks) { return cdiv(nblocks, 64); }

void hd_launch_neighbor_build(hipStream_t st, const HdNeighborArgs& a, bool block_pairs,
                              bool fold_clear) {
  unsigned nblocks = a.npad / 64;
  const bool fold = fold_clear && !a.atom_lists && nblocks <= NB_FOLD_MAX_BLOCKS;
  unsigned nchunks = hd_block_pair_chunks(nblocks);
  unsigned pair_waves = min(nblocks * nchunks, (unsigned)NB_PAIR_WAVES);
  if (block_pairs)
    hipLaunchKernelGGL(k_block_pairs, dim3(cdiv(pair_waves * 64, 256)), dim3(256), 0, st,
                       a.center, a.half, a.lattice, a.rs2, a.block_mask, nblocks, nchunks,
                       a.atom_lists, a.status);
  if (a.atom_lists)
    hipLaunchKernelGGL(k_build_alist, dim3(cdiv(nblocks, NB_WAVES)), dim3(64 * NB_WAVES), 0, st,
                       a.tile_posq, a.tile_key, a.excl_span, a.center, a.half, a.lattice, a.rs2,
                       a.block_mask, nchunks, a.entry_atoms, a.entry_count, a.cap, a.max_count,
                       a.status, a.n, nblocks);
  else if (fold)
    hipLaunchKernelGGL(k_build_entries<true>, dim3(cdiv(nblocks, NB_WAVES)), dim3(64 * NB_WAVES),
                       0, st, a.tile_posq, a.center, a.half, a.lattice, a.rs2, a.block_mask,
                       nchunks, a.entry_atoms, a.entry_count, a.cap, a.max_count, a.status, a.n,
                       nblocks, a.build_ticket, a.rebuild_count, a.block_order);
  else
    hipLaunchKernelGGL(k_build_entries<false>, dim3(cdiv(nblocks, NB_WAVES)), dim3(64 * NB_WAVES),
                       0, st, a.tile_posq, a.center, a.half, a.lattice, a.rs2, a.block_mask,
                       nchunks, a.entry_atoms, a.entry_count, a.cap, a.max_count, a.status, a.n,
                       nblocks, nullptr, nullptr, nullptr);
  if (!fold)
    hipLaunchKernelGGL(k_rebuild_clear, dim3(1), dim3(NB_ORDER_BINS), 0, st, a.status,
                       a.rebuild_count, a.entry_count, nblocks, a.block_order);
}
