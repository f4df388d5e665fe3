// The mesh, from the mesh producer's arrays to the rows the step's kernels read (kernels.hpp), on the device. The input
// is the voronoi-cpp output contract as finite_volume.hpp:257-327 and :353-364 read it, in device memory: owner and
// neighbour per face, area vectors and face centroids (relative to the owner's generator) per face, the CSR lists with
// the owner bit, and per cell the centroid (relative to its generator), volume and position. The reference reads those
// views in place inside its timed kernels (the face offset in its sweep, :322-326; the radius in primitives(), :413);
// here one pass per mesh derives what they derive and packs it:
//   k_geo_cells   per cell: the list length and the geometry half of the cell row [x, c, V, r]; per 64-cell slice (one
//                 work-group) the longest list; checks of the CSR offsets;
//   k_geo_scan    the slices' first slots (exclusive sum of 64 x longest), one work-group;
//   k_geo_pairs   per face: checks of the cell indices (the step reads the face arrays in place);
//   k_geo_slots   per cell and list entry, in the slices' layout: the entry, the other cell, |A| and the face offset
//                 x_f - c_i; padding slots zero; checks of the face indices;
//   k_geo_faces   per face: the record [A, fc, |A|, (owner, neighbour)] the derivatives read (on their first use of a
//                 mesh, not in the pass).
// Every read is guarded by the checks before it, so a malformed mesh sets a flag (read by the host) and reads nothing out
// of range. The arithmetic is that of the host restatement (the same products, sums and rounding), so the rows are equal
// to the bit whichever side derives them.
#include "kernels.hpp"

namespace luna {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ void flag(unsigned int* flags, unsigned int bit) { atomicOr(flags, bit); }

template <int D>
__global__ __launch_bounds__(SLICE) void k_geo_cells(MeshIn m, double* cell, double* volume, int* count, int* longest, unsigned int* flags) {
    using C = CellAt<D>;
    const int i = blockIdx.x * SLICE + threadIdx.x;
    const long long E = 2LL * m.F;
    int n = 0;
    if (i < m.N) {
        const int first = m.start[i], last = m.start[i + 1];
        if (first < 0 || last < first || last > E || (i == 0 && first != 0) || (i == m.N - 1 && last != E)) {
            flag(flags, GEO_BAD_START);
        } else {
            n = last - first;
        }
        count[i] = n;
        double row[GEO_ROW];
#pragma unroll
        for (int k = 0; k < GEO_ROW; ++k) row[k] = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            row[C::X + a] = m.positions[(size_t)i * D + a];
            row[C::C + a] = m.centroid[(size_t)i * D + a];
        }
        const double v = m.volume[i];
        row[C::VOL] = v;
        volume[i] = v;
        row[C::RAD] = cell_radius<D>(v);
        double* out = cell + (size_t)i * CELL_ROW;
#pragma unroll
        for (int k = 0; k < GEO_ROW; ++k) out[k] = row[k];
    }
    // the slice's longest list (one wavefront per work-group)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n = max(n, __shfl_xor(n, o));
    if (threadIdx.x == 0) longest[blockIdx.x] = n;
}

// In place: slice[g] = 64 sum_{h < g} longest[h]; info[0] = the total slots. One work-group of SCAN threads, in tiles of
// SCAN x PER slices: each thread loads PER consecutive slices (the tile's loads coalesced); the threads' sums are scanned
// within each wavefront by shuffles and across the 16 wavefronts through LDS; a carry runs from tile to tile.
constexpr int SCAN = 1024, PER = 16, TILE = SCAN * PER, SCAN_WAVES = SCAN / 64;
__device__ __forceinline__ long long wave_inclusive_sum(long long x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}
__global__ __launch_bounds__(SCAN) void k_geo_scan(int* slice, int S, long long* info, unsigned int* flags) {
    __shared__ long long wave_total[SCAN_WAVES];
    const int t = threadIdx.x, w = t >> 6;
    long long carry = 0;
    for (int base = 0; base < S; base += TILE) {
        const int lo = base + t * PER;
        long long v[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            v[k] = lo + k < S ? (long long)slice[lo + k] * SLICE : 0;
            sum += v[k];
        }
        const long long in_wave = wave_inclusive_sum(sum);
        if ((t & 63) == 63) wave_total[w] = in_wave;
        __syncthreads();
        long long before = 0, tile = 0;  // the earlier wavefronts' totals, and the tile's
#pragma unroll
        for (int k = 0; k < SCAN_WAVES; ++k) {
            const long long x = wave_total[k];
            before += k < w ? x : 0;
            tile += x;
        }
        long long run = carry + before + in_wave - sum;
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (lo + k < S) {
                slice[lo + k] = run > 0x7fffffffLL ? 0 : (int)run;
                run += v[k];
            }
        carry += tile;
        __syncthreads();  // wave_total is rewritten by the next tile
    }
    if (t == 0) {
        info[0] = carry;
        if (carry > 0x7fffffffLL) flag(flags, GEO_TOO_MANY_SLOTS);
    }
}

template <int D>
__global__ __launch_bounds__(BLOCK) void k_geo_faces(MeshIn m, double* face, unsigned int* flags) {
    using Fa = FaceAt<D>;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= m.F) return;
    double row[FACE_ROW];
#pragma unroll
    for (int k = 0; k < FACE_ROW; ++k) row[k] = 0.0;
    double a[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        a[k] = m.area[(size_t)f * D + k];
        row[Fa::A + k] = a[k];
        row[Fa::FC + k] = m.face_centroid[(size_t)f * D + k];
    }
    row[Fa::MAG] = sqrt(dot<D>(a, a));
    const int o = m.owner[f], j = m.neighbour[f];
    if (o < 0 || o >= m.N || j < 0 || j >= m.N) flag(flags, GEO_BAD_CELL);
    int pair[2] = {o, j};
    __builtin_memcpy(&row[Fa::PAIR], pair, sizeof pair);
    double* out = face + (size_t)f * FACE_ROW;
#pragma unroll
    for (int k = 0; k < FACE_ROW; ++k) out[k] = row[k];
}

// The faces' cell indices in range (the flux kernels read owner and neighbour of every face in place).
__global__ __launch_bounds__(BLOCK) void k_geo_pairs(MeshIn m, unsigned int* flags) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= m.F) return;
    const int o = m.owner[f], j = m.neighbour[f];
    if (o < 0 || o >= m.N || j < 0 || j >= m.N) flag(flags, GEO_BAD_CELL);
}

// One work-group per quarter slice: 16 cells, whose list entries are one contiguous run of the CSR list. The work-items
// take consecutive entries (so the list and the faces a cell owns, numbered consecutively, are read in runs); each finds
// its cell among the 16 offsets and computes its slot's values into LDS, laid out as the slots are; then the work-group
// writes each (entry t, array) as one whole run of 16 slots (one 128-byte line for the doubles). Padding slots (t at or
// past the cell's list length) get zeros. A quarter deeper than QUARTER_DEPTH entries writes the rest straight to memory.
// The CSR offsets were checked by k_geo_cells (the host runs this only when they passed).
constexpr int SLOT_BLOCK = 256, QUARTER = 16, QUARTER_DEPTH = 48;
template <int D>
__global__ __launch_bounds__(SLOT_BLOCK) void k_geo_slots(MeshIn m, const int* __restrict__ count, const int* __restrict__ slice,
                                                          const long long* __restrict__ info, int S, int* __restrict__ list,
                                                          int* __restrict__ other, double* __restrict__ slot, unsigned int* flags) {
    __shared__ int offs[QUARTER + 1];
    __shared__ int s_list[QUARTER_DEPTH * QUARTER], s_other[QUARTER_DEPTH * QUARTER];
    __shared__ double s_geo[D + 1][QUARTER_DEPTH * QUARTER];
    const int g = blockIdx.x >> 2, quarter = blockIdx.x & 3;
    const int c0 = g * SLICE + quarter * QUARTER, cells = max(0, min(QUARTER, m.N - c0));
    const int tid = threadIdx.x;
    if (cells > 0)
        for (int l = tid; l <= cells; l += SLOT_BLOCK) offs[l] = m.start[c0 + l];
    const long long slots = info[0];
    const long long base = slice[g];
    const long long end = g + 1 < S ? (long long)slice[g + 1] : slots;
    const int depth = (int)((end - base) / SLICE);
    const int staged = min(depth, QUARTER_DEPTH);
    for (int p = tid; p < staged * QUARTER; p += SLOT_BLOCK) {
        s_list[p] = 0;
        s_other[p] = 0;
#pragma unroll
        for (int a = 0; a <= D; ++a) s_geo[a][p] = 0.0;
    }
    __syncthreads();
    const int k0 = cells > 0 ? offs[0] : 0, k1 = cells > 0 ? offs[cells] : 0;
    const int* __restrict__ in_list = m.list;
    const double* __restrict__ pos = m.positions;
    const double* __restrict__ cen = m.centroid;
    unsigned int bad = 0;
    for (int k = k0 + tid; k < k1; k += SLOT_BLOCK) {
        int lane = 0;  // the last l with offs[l] <= k
#pragma unroll
        for (int step = QUARTER / 2; step > 0; step >>= 1)
            if (lane + step < cells && offs[lane + step] <= k) lane += step;
        const int i = c0 + lane, t = k - offs[lane];
        int e = in_list[k];
        const int f = e >= 0 ? e : ~e;
        int cell = 0;
        double mag = 0.0, off[D];
#pragma unroll
        for (int a = 0; a < D; ++a) off[a] = 0.0;
        if (f >= m.F) {
            bad |= GEO_BAD_FACE;
            e = 0;
        } else {
            const int o = m.owner[f], j = m.neighbour[f];
            if (o < 0 || o >= m.N || j < 0 || j >= m.N) {
                bad |= GEO_BAD_CELL;
            } else if (o != i && j != i) {
                bad |= GEO_BAD_FACE;  // a face of other cells in this cell's list
            } else {
                // The owner bit from the face's owner (finite_volume.hpp:78-87 owner_layout: set when the mesh's lists
                // lack it; a list that carries it keeps its bits).
                e = o == i ? f : ~f;
                cell = e >= 0 ? j : o;
                double A[D], rel[D];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    A[a] = m.area[(size_t)f * D + a];
                    rel[a] = m.face_centroid[(size_t)f * D + a];
                }
                mag = sqrt(dot<D>(A, A));  // as k_geo_faces and the flux kernels take it
                // Geometry::face_offset (finite_volume.hpp:322-326): the centroid moved to this cell's generator by the
                // nearest-image separation when the cell does not own the face, then relative to its centroid
                if (e < 0) {
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        const double d = pos[(size_t)o * D + a] - pos[(size_t)i * D + a];
                        rel[a] = rel[a] + (d - m.side[a] * round(d * m.inverse[a]));
                    }
                }
#pragma unroll
                for (int a = 0; a < D; ++a) off[a] = rel[a] - cen[(size_t)i * D + a];
            }
        }
        if (t < QUARTER_DEPTH) {
            const int p = t * QUARTER + lane;
            s_list[p] = e;
            s_other[p] = cell;
            s_geo[0][p] = mag;
#pragma unroll
            for (int a = 0; a < D; ++a) s_geo[1 + a][p] = off[a];
        } else {  // beyond the staging: straight to memory
            const long long sl = base + (long long)t * SLICE + quarter * QUARTER + lane;
            list[sl] = e;
            other[sl] = cell;
            slot[sl] = mag;
#pragma unroll
            for (int a = 0; a < D; ++a) slot[(a + 1) * slots + sl] = off[a];
        }
    }
    __syncthreads();
    // the staged slots, a run of 16 per entry and array; then the padding past the staging
    for (int p = tid; p < staged * QUARTER; p += SLOT_BLOCK) {
        const int t = p / QUARTER, l = p % QUARTER;
        const long long sl = base + (long long)t * SLICE + quarter * QUARTER + l;
        list[sl] = s_list[p];
        other[sl] = s_other[p];
        slot[sl] = s_geo[0][p];
#pragma unroll
        for (int a = 0; a < D; ++a) slot[(a + 1) * slots + sl] = s_geo[1 + a][p];
    }
    for (int p = staged * QUARTER + tid; p < depth * QUARTER; p += SLOT_BLOCK) {
        const int t = p / QUARTER, l = p % QUARTER;
        const int n = l < cells ? offs[l + 1] - offs[l] : 0;
        if (t < n) continue;
        const long long sl = base + (long long)t * SLICE + quarter * QUARTER + l;
        list[sl] = 0;
        other[sl] = 0;
    }
    if (bad) flag(flags, bad);
}

}  // namespace

template <int D>
hipError_t launch_geometry_cells(hipStream_t s, const MeshIn& m, double* cell, double* volume, int* count, int* slice, unsigned int* flags) {
    const int S = (m.N + SLICE - 1) / SLICE;
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geo_cells<D>, dim3(S), dim3(SLICE), 0, s, m, cell, volume, count, slice, flags);
    return hipGetLastError();
}

hipError_t launch_geometry_scan(hipStream_t s, int* slice, int S, long long* info, unsigned int* flags) {
    hipLaunchKernelGGL(k_geo_scan, dim3(1), dim3(SCAN), 0, s, slice, S, info, flags);
    return hipGetLastError();
}

template <int D>
hipError_t launch_geometry_faces(hipStream_t s, const MeshIn& m, double* face, unsigned int* flags) {
    if (m.F == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geo_faces<D>, dim3((m.F + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, m, face, flags);
    return hipGetLastError();
}

hipError_t launch_geometry_pairs(hipStream_t s, const MeshIn& m, unsigned int* flags) {
    if (m.F == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geo_pairs, dim3((m.F + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, m, flags);
    return hipGetLastError();
}

template <int D>
hipError_t launch_geometry_slots(hipStream_t s, const MeshIn& m, const int* count, const int* slice, const long long* info, int* list,
                                 int* other, double* slot, unsigned int* flags) {
    const int S = (m.N + SLICE - 1) / SLICE;
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_geo_slots<D>, dim3(S * (SLICE / QUARTER)), dim3(SLOT_BLOCK), 0, s, m, count, slice, info, S, list, other, slot, flags);
    return hipGetLastError();
}

template hipError_t launch_geometry_cells<2>(hipStream_t, const MeshIn&, double*, double*, int*, int*, unsigned int*);
template hipError_t launch_geometry_cells<3>(hipStream_t, const MeshIn&, double*, double*, int*, int*, unsigned int*);
template hipError_t launch_geometry_faces<2>(hipStream_t, const MeshIn&, double*, unsigned int*);
template hipError_t launch_geometry_faces<3>(hipStream_t, const MeshIn&, double*, unsigned int*);
template hipError_t launch_geometry_slots<2>(hipStream_t, const MeshIn&, const int*, const int*, const long long*, int*, int*, double*, unsigned int*);
template hipError_t launch_geometry_slots<3>(hipStream_t, const MeshIn&, const int*, const int*, const long long*, int*, int*, double*, unsigned int*);

}  // namespace luna
