// gfx950 kernels of the MPM snow substep.
//
// Layout in HBM (one scene):
//   particles  field-major FP32 arrays (X[D], V[D], F[D*D], Jp, m, V0) of
//              capacity pcap, double buffered; after binning they are
//              grouped by home block so every block's particles are one
//              contiguous range.
//   grid       block-sparse: the padded domain is cut into 4^3 (3D) or
//              8^2 (2D) node blocks; a dense slot table maps block
//              coordinates to a compacted ordinal; every channel is an
//              array of ordinal*64 + local-node entries.
// Block kernels (P2G, force, CG force, G2P) run one work-group per active
// block: the block's particles touch only its nodes and a one-node halo,
// a (B+2)^D tile that lives in LDS.  Scatters accumulate with LDS float
// atomics and flush the tile once with global float atomics; gathers load
// the tile once from the 3^D neighbour blocks.
#include "mpm_kernels.h"

namespace mpm {

template <int D>
struct PF {
    static constexpr int X = 0, V = D, F = 2 * D, J = 2 * D + D * D, M = J + 1, VOL = J + 2, N = J + 3;
};

// Work-groups are dispatched round-robin over the 8 XCDs; map them so
// each XCD walks a contiguous range of blocks (neighbouring blocks share
// halo nodes in that XCD's L2).  Requires gridDim.x % 8 == 0.
__device__ inline int xcd_wg() {
    const int G = (int)gridDim.x, w = (int)blockIdx.x;
    return (w & 7) * (G >> 3) + (w >> 3);
}

template <int D>
__device__ inline void slot_coords(int slot, int nb, int* b) {
    if constexpr (D == 3) {
        b[2] = slot % nb;
        b[1] = (slot / nb) % nb;
        b[0] = slot / (nb * nb);
    } else {
        b[1] = slot % nb;
        b[0] = slot / nb;
    }
}

// tile index -> (neighbour ordinal in 0..3^D-1, local node in 0..63, tile coords)
template <int D>
__device__ inline void tile_map(int t, int& nbr, int& local, int* tc) {
    using Dm = Dims<D>;
    constexpr int TE = Dm::TE, B = Dm::B;
    if constexpr (D == 3) {
        tc[0] = t / (TE * TE);
        tc[1] = (t / TE) % TE;
        tc[2] = t % TE;
    } else {
        tc[0] = t / TE;
        tc[1] = t % TE;
    }
    nbr = 0;
    local = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const int o = tc[a] == 0 ? 0 : (tc[a] == TE - 1 ? 2 : 1);
        const int l = tc[a] - 1 - B * (o - 1);
        nbr = nbr * 3 + o;
        local = local * B + l;
    }
}

// stencil offsets of node s (last axis fastest) -> tile index relative to the local cell
template <int D>
__device__ inline int tile_index(const int* l, int s, int* off) {
    using Dm = Dims<D>;
    constexpr int TE = Dm::TE;
    if constexpr (D == 3) {
        off[0] = s / 9;
        off[1] = (s / 3) % 3;
        off[2] = s % 3;
        return ((l[0] + off[0]) * TE + (l[1] + off[1])) * TE + (l[2] + off[2]);
    } else {
        off[0] = s / 3;
        off[1] = s % 3;
        return (l[0] + off[0]) * TE + (l[1] + off[1]);
    }
}

// Weight (Tools/Influence_Iterator.h:85-90): left-to-right product over axes.
template <int D>
__device__ inline float stencil_weight(const float (*w)[3], const int* off) {
#pragma clang fp contract(off)
    float r = w[0][off[0]];
#pragma unroll
    for (int a = 1; a < D; ++a) r = r * w[a][off[a]];
    return r;
}
// Weight_Gradient (Tools/Influence_Iterator.h:92-99)
template <int D>
__device__ inline void stencil_grad(const float (*w)[3], const float (*dw)[3], const int* off, float* g) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < D; ++i) {
        float r = (i == 0) ? dw[0][off[0]] : w[0][off[0]];
#pragma unroll
        for (int j = 1; j < D; ++j) r = r * ((i == j) ? dw[j][off[j]] : w[j][off[j]]);
        g[i] = r;
    }
}

// Weights of particle i and its cell relative to block bb.  Returns false
// (and the caller skips the particle) when the cell is not inside the block,
// which can only happen if particles moved after the last binning; no tile
// index outside [0, B+1]^D is ever formed.
template <int D>
__device__ inline bool particle_weights(const KP& p, int i, const int* bb, float (*w)[3], float (*dw)[3], int* l,
                                        float* X) {
    constexpr int B = Dims<D>::B;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        X[a] = p.pcur[(PF<D>::X + a) * p.pcap + i];
        int c;
        bspline_weights(X[a], p.dx, p.inv_dx, c, w[a], dw[a]);
        l[a] = c + B - B * bb[a];
        ok = ok && l[a] >= 0 && l[a] < B;
    }
    return ok;
}

// x crop of the reference's slab decomposition: scatters never reach nodes
// whose x index is outside [0, counts-1].
template <int D>
__device__ inline bool flush_allowed(const KP& p, const int* bb, const int* tc) {
    if (!p.crop_ghost_x) return true;
    constexpr int B = Dims<D>::B;
    const int ix = B * bb[0] - 1 + tc[0] - B;
    return ix >= 0 && ix <= p.counts - 1;
}

__device__ inline void wg_reduce2(double& sum, double& mx) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        mx = fmax(mx, __shfl_xor(mx, off));
    }
    __shared__ double ss[WG / 64], sm[WG / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) { ss[wave] = sum; sm[wave] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, m = 0.0;
#pragma unroll
        for (int k = 0; k < WG / 64; ++k) { s += ss[k]; m = fmax(m, sm[k]); }
        sum = s;
        mx = m;
    }
}

// Publishes the work-group's partials and returns true in the last
// work-group to arrive; that group then sees every partial.
__device__ inline bool publish_and_check_last(const KP& p, double sum, double mx) {
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        p.red[blockIdx.x] = sum;
        p.red[gridDim.x + blockIdx.x] = mx;
        __threadfence();
        const unsigned prev = __hip_atomic_fetch_add(p.sync, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (is_last) __threadfence();
    return is_last != 0;
}

// Sums the partials in a fixed order (deterministic) inside the last group.
__device__ inline void gather_partials(const KP& p, double& sum, double& mx) {
    double s = 0.0, m = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += WG) {
        s += __hip_atomic_load(&p.red[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        m = fmax(m, __hip_atomic_load(&p.red[gridDim.x + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    wg_reduce2(s, m);
    sum = s;
    mx = m;
}

__device__ inline bool inside_plane(const KP& p, int id, const float* x, int D) {
    float s = 0.0f;
    for (int a = 0; a < D; ++a) s += (x[a] - p.ps[id][a]) * p.pn[id][a];
    return s < 0.0f;
}

template <int D>
__device__ inline void node_center(const KP& p, int n, float* x) {
    constexpr int B = Dims<D>::B;
    const int b = n >> 6, l = n & 63;
    int bb[D];
    slot_coords<D>(p.block_slot[b], p.nb, bb);
    int lc[D];
    if constexpr (D == 3) { lc[0] = l >> 4; lc[1] = (l >> 2) & 3; lc[2] = l & 3; }
    else { lc[0] = l >> 3; lc[1] = l & 7; }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const int i = B * bb[a] + lc[a] - B;
        x[a] = ((float)i + 0.5f) * p.dx;
    }
}

// ============================================================================
// Binning: key = home block slot, rank within the block, activation marks
// ============================================================================
template <int D>
__global__ __launch_bounds__(WG) void k_bin(KP p) {
    constexpr int B = Dims<D>::B;
    const int n = p.ctr->n_live;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        if (p.keycur[i] < 0) continue;
        int c[D];
        bool ok = true, finite = true;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const float X = p.pcur[(PF<D>::X + a) * p.pcap + i];
            finite = finite && isfinite(X);
            c[a] = finite ? cell_of(X, p.inv_dx) : -1;
            ok = ok && finite && c[a] >= 0 && c[a] <= p.counts - 1;
        }
        if (!ok) {
            p.keycur[i] = -1;
            if (!finite) atomicOr(&p.ctr->nonfinite, 1);
            continue;
        }
        int slot = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) slot = slot * p.nb + (c[a] + B) / B;
        p.prank[i] = atomicAdd(&p.slot_count[slot], 1);
        p.keycur[i] = slot;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < D; ++a) { lo[a] = (c[a] + B - 1) / B; hi[a] = (c[a] + B + 1) / B; }
        if constexpr (D == 3) {
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) p.slot_touched[(x * p.nb + y) * p.nb + z] = 1;
        } else {
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y) p.slot_touched[x * p.nb + y] = 1;
        }
    }
}

// exclusive scan of one int per thread over the work-group
__device__ inline int wg_exclusive_scan(int v, int& total) {
    __shared__ int wsum[WG / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < WG / 64; ++k) {
        if (k < wave) before += wsum[k];
        total += wsum[k];
    }
    return before + x - v;
}

__global__ __launch_bounds__(WG) void k_scan_reduce(KP p) {
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    int act = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        const int s = base + k;
        if (s < p.nslots) {
            act += p.slot_touched[s] != 0;
            cnt += p.slot_count[s];
        }
    }
    int ta, tc;
    wg_exclusive_scan(act, ta);
    wg_exclusive_scan(cnt, tc);
    if (threadIdx.x == 0) {
        p.scan_part[blockIdx.x] = ta;
        p.scan_part[p.nparts + blockIdx.x] = tc;
    }
}

__global__ __launch_bounds__(WG) void k_scan_top(KP p) {
    // one work-group: exclusive scan of nparts block totals (two arrays)
    const int per = (p.nparts + WG - 1) / WG;
    const int b0 = threadIdx.x * per;
    int sa = 0, sc = 0;
    for (int k = 0; k < per; ++k) {
        const int j = b0 + k;
        if (j < p.nparts) { sa += p.scan_part[j]; sc += p.scan_part[p.nparts + j]; }
    }
    int ta, tc;
    const int ea = wg_exclusive_scan(sa, ta);
    const int ec = wg_exclusive_scan(sc, tc);
    int ra = ea, rc = ec;
    for (int k = 0; k < per; ++k) {
        const int j = b0 + k;
        if (j < p.nparts) {
            const int va = p.scan_part[j], vc = p.scan_part[p.nparts + j];
            p.scan_part[j] = ra;
            p.scan_part[p.nparts + j] = rc;
            ra += va;
            rc += vc;
        }
    }
    if (threadIdx.x == 0) {
        p.ctr->n_prev = p.ctr->n_live;
        p.ctr->n_live = tc;
        p.ctr->nblocks = ta;
        p.ctr->max_v2 = 0u;
        p.ctr->n_removed = 0;
        if (ta > p.bcap) p.ctr->overflow = 1;
        p.block_pstart[ta < p.bcap ? ta : p.bcap] = tc;
    }
}

__global__ __launch_bounds__(WG) void k_scan_apply(KP p) {
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    int act[SCAN_PER_THREAD], cnt[SCAN_PER_THREAD];
    int sa = 0, sc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        const int s = base + k;
        act[k] = (s < p.nslots) ? (p.slot_touched[s] != 0) : 0;
        cnt[k] = (s < p.nslots) ? p.slot_count[s] : 0;
        sa += act[k];
        sc += cnt[k];
    }
    int ta, tc;
    int ra = wg_exclusive_scan(sa, ta) + p.scan_part[blockIdx.x];
    int rc = wg_exclusive_scan(sc, tc) + p.scan_part[p.nparts + blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        const int s = base + k;
        if (s >= p.nslots) break;
        if (act[k]) {
            p.slot_active[s] = ra;
            if (ra < p.bcap) {
                p.block_slot[ra] = s;
                p.block_pstart[ra] = rc;
            }
        } else {
            p.slot_active[s] = -1;
        }
        p.slot_touched[s] = 0;
        p.slot_count[s] = 0;
        ra += act[k];
        rc += cnt[k];
    }
}

// neighbour table + clearing of the scatter channels of every active block
template <int D>
__global__ __launch_bounds__(WG) void k_block_setup(KP p) {
    using Dm = Dims<D>;
    if (p.ctr->overflow) return;
    const int nblocks = p.ctr->nblocks;
    const int group = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = blockIdx.x * (WG / 64) + group; b < nblocks; b += gridDim.x * (WG / 64)) {
        if (lane < Dm::NBR) {
            int bb[D];
            slot_coords<D>(p.block_slot[b], p.nb, bb);
            int o[3];
            if constexpr (D == 3) { o[0] = lane / 9; o[1] = (lane / 3) % 3; o[2] = lane % 3; }
            else { o[0] = lane / 3; o[1] = lane % 3; }
            int s = 0;
            bool in = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const int c = bb[a] + o[a] - 1;
                in = in && c >= 0 && c < p.nb;
                s = s * p.nb + c;
            }
            p.block_nbr[b * Dm::NBR + lane] = in ? p.slot_active[s] : -1;
        }
        const int64_t n = (int64_t)b * 64 + lane;
        p.g_m[n] = 0.0f;
        p.g_flags[n] = 0u;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            p.g_v[a * p.gcap + n] = 0.0f;
            p.g_q[a * p.gcap + n] = 0.0f;
        }
    }
}

template <int D>
__global__ __launch_bounds__(WG) void k_reorder(KP p) {
    if (p.ctr->overflow) return;
    const int n = p.ctr->n_prev;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const int key = p.keycur[i];
        if (key < 0) continue;
        const int dest = p.block_pstart[p.slot_active[key]] + p.prank[i];
#pragma unroll
        for (int f = 0; f < PF<D>::N; ++f) p.palt[f * p.pcap + dest] = p.pcur[f * p.pcap + i];
        p.idalt[dest] = p.idcur[i];
        p.keyalt[dest] = 0;
    }
}

// ============================================================================
// P2G (Rasterize, MPM_Example.cpp:520-537): m_i += w m_p, (mv)_i += w m_p V_p
// ============================================================================
template <int D>
__global__ __launch_bounds__(WG) void k_p2g(KP p) {
    using Dm = Dims<D>;
    constexpr int TN = Dm::TN, NBR = Dm::NBR;
    __shared__ float tile[D + 1][TN];
    __shared__ int nbr[NBR];
    if (p.ctr->overflow) return;
    const int nblocks = p.ctr->nblocks;
    for (int b = xcd_wg(); b < nblocks; b += gridDim.x) {
        for (int t = threadIdx.x; t < (D + 1) * TN; t += WG) (&tile[0][0])[t] = 0.0f;
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        slot_coords<D>(p.block_slot[b], p.nb, bb);
        __syncthreads();
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + threadIdx.x; i < i1; i += WG) {
            float w[D][3], dw[D][3], X[D];
            int l[D];
            if (!particle_weights<D>(p, i, bb, w, dw, l, X)) continue;
            const float m = p.pcur[PF<D>::M * p.pcap + i];
            float mv[D];
#pragma unroll
            for (int a = 0; a < D; ++a) mv[a] = m * p.pcur[(PF<D>::V + a) * p.pcap + i];
#pragma unroll
            for (int s = 0; s < Dm::ST; ++s) {
                int off[D];
                const int t = tile_index<D>(l, s, off);
                const float wt = stencil_weight<D>(w, off);
                atomicAdd(&tile[0][t], wt * m);
#pragma unroll
                for (int a = 0; a < D; ++a) atomicAdd(&tile[1 + a][t], wt * mv[a]);
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < TN; t += WG) {
            const float mm = tile[0][t];
            if (mm == 0.0f) continue;
            int nb, local, tc[D];
            tile_map<D>(t, nb, local, tc);
            const int blk = nbr[nb];
            if (blk < 0 || !flush_allowed<D>(p, bb, tc)) continue;
            const int64_t node = (int64_t)blk * 64 + local;
            unsafeAtomicAdd(&p.g_m[node], mm);
#pragma unroll
            for (int a = 0; a < D; ++a) unsafeAtomicAdd(&p.g_v[a * p.gcap + node], tile[1 + a][t]);
        }
        __syncthreads();
    }
}

// Flag_Setup_Helper (Flag_Setup_Helper.h:25-35) + Velocity_Normalization_Helper
// (Velocity_Normalization_Helper.h:39-52)
template <int D>
__global__ __launch_bounds__(WG) void k_normalize(KP p) {
    if (p.ctr->overflow) return;
    const int nn = p.ctr->nblocks * 64;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        const float m = p.g_m[n];
        const bool interior = m > 0.0f;
        p.g_flags[n] = interior ? MPM_FLAG_INTERIOR_DEV : 0u;
        if (interior) {
            const float inv = 1.0f / m;
#pragma unroll
            for (int a = 0; a < D; ++a) p.g_v[a * p.gcap + n] *= inv;
        }
    }
}

// gather of `nch` grid channel arrays into an LDS tile
template <int D, int NCH>
__device__ inline void gather_tile(const KP& p, const int* nbr, float (*tile)[Dims<D>::TN], const float* const* src) {
    constexpr int TN = Dims<D>::TN;
    for (int t = threadIdx.x; t < TN; t += WG) {
        int nb, local, tc[D];
        tile_map<D>(t, nb, local, tc);
        const int blk = nbr[nb];
#pragma unroll
        for (int c = 0; c < NCH; ++c) tile[c][t] = blk >= 0 ? src[c][(int64_t)blk * 64 + local] : 0.0f;
    }
}

// Process_Waiting_Particles (MPM_Example.cpp:827-840): V0 = m / (sum w m_i / dx^d)
template <int D>
__global__ __launch_bounds__(WG) void k_init_volume(KP p) {
    using Dm = Dims<D>;
    constexpr int TN = Dm::TN, NBR = Dm::NBR;
    __shared__ float tile[1][TN];
    __shared__ int nbr[NBR];
    if (p.ctr->overflow) return;
    const int nblocks = p.ctr->nblocks;
    float vol = 1.0f;
    for (int a = 0; a < D; ++a) vol *= p.dx;
    const float inv_vol = 1.0f / vol;
    for (int b = xcd_wg(); b < nblocks; b += gridDim.x) {
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        slot_coords<D>(p.block_slot[b], p.nb, bb);
        __syncthreads();
        const float* src[1] = {p.g_m};
        gather_tile<D, 1>(p, nbr, tile, src);
        __syncthreads();
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + threadIdx.x; i < i1; i += WG) {
            float w[D][3], dw[D][3], X[D];
            int l[D];
            if (!particle_weights<D>(p, i, bb, w, dw, l, X)) continue;
            float rho = 0.0f;
#pragma unroll
            for (int s = 0; s < Dm::ST; ++s) {
                int off[D];
                const int t = tile_index<D>(l, s, off);
                rho += stencil_weight<D>(w, off) * tile[0][t];
            }
            rho *= inv_vol;
            p.pcur[PF<D>::VOL * p.pcap + i] = p.pcur[PF<D>::M * p.pcap + i] / rho;
        }
        __syncthreads();
    }
}

// ============================================================================
// Constitutive update (MPM_Example.cpp:563-571) and explicit force P2G
// (MPM_Example.cpp:757-784): f_i += -V0 P F^T grad w + g m w
// ============================================================================
template <int D, bool CONST, bool FORCE>
__global__ __launch_bounds__(WG) void k_particle_force(KP p) {
    using Dm = Dims<D>;
    constexpr int TN = Dm::TN, NBR = Dm::NBR, DD = D * D;
    __shared__ float tile[D][TN];
    __shared__ int nbr[NBR];
    if (p.ctr->overflow) return;
    const int nblocks = p.ctr->nblocks;
    for (int b = xcd_wg(); b < nblocks; b += gridDim.x) {
        int bb[D];
        slot_coords<D>(p.block_slot[b], p.nb, bb);
        if (FORCE) {
            for (int t = threadIdx.x; t < D * TN; t += WG) (&tile[0][0])[t] = 0.0f;
            if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
            __syncthreads();
        }
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + threadIdx.x; i < i1; i += WG) {
            float F[DD], R[DD];
#pragma unroll
            for (int k = 0; k < DD; ++k) F[k] = p.pcur[(PF<D>::F + k) * p.pcap + i];
            float Jp = p.pcur[PF<D>::J * p.pcap + i];
            if (CONST) {
                plasticity<D>(F, Jp, R, p.mat);
#pragma unroll
                for (int k = 0; k < DD; ++k) {
                    p.pcur[(PF<D>::F + k) * p.pcap + i] = F[k];
                    p.rcache[k * p.pcap + i] = R[k];
                }
                p.pcur[PF<D>::J * p.pcap + i] = Jp;
            } else {
#pragma unroll
                for (int k = 0; k < DD; ++k) R[k] = p.rcache[k * p.pcap + i];
            }
            if (FORCE) {
                StressState<D> st;
                stress_state<D>(F, R, Jp, p.mat, st);
                float P[DD];
                first_pk<D>(st, P);
                const float V0 = p.pcur[PF<D>::VOL * p.pcap + i];
                const float m = p.pcur[PF<D>::M * p.pcap + i];
                float A[DD];  // V0 P F^T
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        float s = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) s += P[r * D + k] * F[c * D + k];
                        A[r * D + c] = s * V0;
                    }
                float w[D][3], dw[D][3], X[D];
                int l[D];
                if (!particle_weights<D>(p, i, bb, w, dw, l, X)) continue;
                float gm[D];
#pragma unroll
                for (int a = 0; a < D; ++a) gm[a] = p.gravity[a] * m;
#pragma unroll
                for (int s = 0; s < Dm::ST; ++s) {
                    int off[D];
                    const int t = tile_index<D>(l, s, off);
                    float g[D];
                    stencil_grad<D>(w, dw, off, g);
                    const float wt = stencil_weight<D>(w, off);
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        float inner = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) inner += A[a * D + k] * g[k];
                        atomicAdd(&tile[a][t], gm[a] * wt - inner);
                    }
                }
            }
        }
        if (FORCE) {
            __syncthreads();
            for (int t = threadIdx.x; t < TN; t += WG) {
                int nb, local, tc[D];
                tile_map<D>(t, nb, local, tc);
                const int blk = nbr[nb];
                if (blk < 0 || !flush_allowed<D>(p, bb, tc)) continue;
                const int64_t node = (int64_t)blk * 64 + local;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const float v = tile[a][t];
                    if (v != 0.0f) unsafeAtomicAdd(&p.g_q[a * p.gcap + node], v);
                }
            }
            __syncthreads();
        }
    }
}

// ============================================================================
// Grid update: Explicit_Force_Helper (v* = v + dt f / m), plane collision
// (Grid_Based_Collision_Helper.h:57-85) and the CG set-up (MPM_RHS_Helper,
// Reset_Solver_Channels: r = b = v*, s = q = 0).
// ============================================================================
template <int D, bool FORCE, bool COLL, bool CGINIT>
__global__ __launch_bounds__(WG) void k_grid_update(KP p) {
    if (p.ctr->overflow) return;
    if (CGINIT && blockIdx.x == 0 && threadIdx.x == 0) {
        p.cg->iter = 0;
        p.cg->done = 0;
        p.cg->converged = 0;
        p.cg->multiplies = 0;
        p.cg->rho = p.cg->rho_old = p.cg->alpha = p.cg->beta = p.cg->conv = p.cg->rr = p.cg->sq = 0.0;
    }
    const int nn = p.ctr->nblocks * 64;
    const float dt = p.dt;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float vs[D];
        if (FORCE) {
            if (interior) {
                const float dtm = dt / p.g_m[n];
#pragma unroll
                for (int a = 0; a < D; ++a) vs[a] = p.g_v[a * p.gcap + n] + dtm * p.g_q[a * p.gcap + n];
            } else {
#pragma unroll
                for (int a = 0; a < D; ++a) vs[a] = 0.0f;
            }
        } else {
#pragma unroll
            for (int a = 0; a < D; ++a) vs[a] = p.g_vs[a * p.gcap + n];
        }
        if (COLL && interior) {
            float x[D];
            node_center<D>(p, n, x);
            for (int id = 0; id < p.nplanes; ++id) {
                if (!inside_plane(p, id, x, D)) continue;
                float proj = 0.0f;
#pragma unroll
                for (int a = 0; a < D; ++a) proj += vs[a] * p.pn[id][a];
                if (proj < 0.0f) {
#pragma unroll
                    for (int a = 0; a < D; ++a) vs[a] -= p.pn[id][a] * proj;
                    float nrm = 0.0f;
#pragma unroll
                    for (int a = 0; a < D; ++a) nrm += vs[a] * vs[a];
                    nrm = sqrtf(nrm);
                    const float mu = p.pmu[id];
                    if (-proj * mu < nrm) {
#pragma unroll
                        for (int a = 0; a < D; ++a) vs[a] += vs[a] / nrm * proj * mu;
                    } else {
#pragma unroll
                        for (int a = 0; a < D; ++a) vs[a] = 0.0f;
                    }
                }
            }
        }
        if (FORCE || COLL) {
#pragma unroll
            for (int a = 0; a < D; ++a) p.g_vs[a * p.gcap + n] = vs[a];
        }
        if (CGINIT) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                p.g_r[a * p.gcap + n] = interior ? vs[a] : 0.0f;
                p.g_s[a * p.gcap + n] = 0.0f;
                p.g_q[a * p.gcap + n] = 0.0f;
            }
        }
    }
}

// ============================================================================
// CG operator: MPM_CG_System::Force (MPM_CG_System.h:150-215)
//   T = sum_i grad w_ip (x) u_i;  dF = T^T F;  A = V0 F dP(dF)^T;  q_i += A^T grad w_ip
// ============================================================================
template <int D, bool RESTART>
__global__ __launch_bounds__(WG) void k_cg_force(KP p) {
    using Dm = Dims<D>;
    constexpr int TN = Dm::TN, NBR = Dm::NBR, DD = D * D;
    __shared__ float ut[D][TN];
    __shared__ float qt[D][TN];
    __shared__ int nbr[NBR];
    if (p.ctr->overflow || p.cg->done) return;
    const int nblocks = p.ctr->nblocks;
    const float* u = RESTART ? p.g_vs : p.g_s;
    for (int b = xcd_wg(); b < nblocks; b += gridDim.x) {
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        slot_coords<D>(p.block_slot[b], p.nb, bb);
        for (int t = threadIdx.x; t < D * TN; t += WG) (&qt[0][0])[t] = 0.0f;
        __syncthreads();
        const float* src[D];
#pragma unroll
        for (int a = 0; a < D; ++a) src[a] = u + a * p.gcap;
        gather_tile<D, D>(p, nbr, ut, src);
        __syncthreads();
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + threadIdx.x; i < i1; i += WG) {
            float w[D][3], dw[D][3], X[D];
            int l[D];
            if (!particle_weights<D>(p, i, bb, w, dw, l, X)) continue;
            float T[DD];
#pragma unroll
            for (int k = 0; k < DD; ++k) T[k] = 0.0f;
#pragma unroll
            for (int s = 0; s < Dm::ST; ++s) {
                int off[D];
                const int t = tile_index<D>(l, s, off);
                float g[D];
                stencil_grad<D>(w, dw, off, g);
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int c = 0; c < D; ++c) T[a * D + c] += g[a] * ut[c][t];
            }
            float F[DD], R[DD];
#pragma unroll
            for (int k = 0; k < DD; ++k) {
                F[k] = p.pcur[(PF<D>::F + k) * p.pcap + i];
                R[k] = p.rcache[k * p.pcap + i];
            }
            const float Jp = p.pcur[PF<D>::J * p.pcap + i];
            const float V0 = p.pcur[PF<D>::VOL * p.pcap + i];
            float dF[DD];
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    float s = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) s += T[k * D + r] * F[k * D + c];
                    dF[r * D + c] = s;
                }
            StressState<D> st;
            stress_state<D>(F, R, Jp, p.mat, st);
            float dP[DD];
            dP_of<D>(st, dF, dP);
            float A[DD];  // V0 F dP^T
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    float s = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) s += F[r * D + k] * dP[c * D + k];
                    A[r * D + c] = s * V0;
                }
#pragma unroll
            for (int s = 0; s < Dm::ST; ++s) {
                int off[D];
                const int t = tile_index<D>(l, s, off);
                float g[D];
                stencil_grad<D>(w, dw, off, g);
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    float v = 0.0f;
#pragma unroll
                    for (int r = 0; r < D; ++r) v += A[r * D + c] * g[r];
                    atomicAdd(&qt[c][t], v);
                }
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < TN; t += WG) {
            int nb, local, tc[D];
            tile_map<D>(t, nb, local, tc);
            const int blk = nbr[nb];
            if (blk < 0 || !flush_allowed<D>(p, bb, tc)) continue;
            const int64_t node = (int64_t)blk * 64 + local;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float v = qt[a][t];
                if (v != 0.0f) unsafeAtomicAdd(&p.g_q[a * p.gcap + node], v);
            }
        }
        __syncthreads();
    }
}

// Project (Project_Helper.h:30-45): zero at nodes inside a plane
template <int D>
__device__ inline bool project_zero(const KP& p, int n) {
    if (p.project_mode != 0 || p.nplanes == 0) return false;
    float x[D];
    node_center<D>(p, n, x);
    for (int id = 0; id < p.nplanes; ++id)
        if (inside_plane(p, id, x, D)) return true;
    return false;
}

// CG restart: q = dt^2/m q + x (Multiply_Helper), r = b - q, Project(r);
// then the convergence norm and r.r.
template <int D>
__global__ __launch_bounds__(WG) void k_cg_restart_node(KP p) {
    if (p.ctr->overflow) return;
    const int nn = p.ctr->nblocks * 64;
    const float sds = p.dt * p.dt;
    double sum = 0.0, mx = 0.0;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float r[D];
#pragma unroll
        for (int a = 0; a < D; ++a) r[a] = p.g_r[a * p.gcap + n];
        if (interior) {
            const float c = sds / p.g_m[n];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float q = c * p.g_q[a * p.gcap + n] + p.g_vs[a * p.gcap + n];
                r[a] = r[a] - q;
            }
        }
        if (project_zero<D>(p, n)) {
#pragma unroll
            for (int a = 0; a < D; ++a) r[a] = 0.0f;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            p.g_r[a * p.gcap + n] = r[a];
            p.g_q[a * p.gcap + n] = 0.0f;
        }
        if (interior) {
            const double m = (double)p.g_m[n];
            double rr = 0.0, r2 = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                rr += (double)r[a] * ((double)r[a] * m);
                r2 += (double)r[a] * (double)r[a];
            }
            sum += rr;
            mx = fmax(mx, r2);
        }
    }
    wg_reduce2(sum, mx);
    if (publish_and_check_last(p, sum, mx)) {
        gather_partials(p, sum, mx);
        if (threadIdx.x == 0) {
            CGState* cg = p.cg;
            cg->multiplies = 1;
            cg->iter = 0;
            cg->conv = sqrt(mx);
            cg->rr = sum;
            if (cg->conv <= p.cg_tol && (0 >= p.cg_min || cg->conv < (double)FLT_EPSILON)) {
                cg->done = 1;
                cg->converged = 1;
            } else if (p.cg_max == 0) {
                cg->done = 1;
            } else {
                cg->rho = sum;
            }
            *p.sync = 0u;
        }
    }
}

// s = r (first iteration) or s = (rho/rho_old) s + r
template <int D>
__global__ __launch_bounds__(WG) void k_cg_update_s(KP p) {
    if (p.ctr->overflow || p.cg->done) return;
    const int nn = p.ctr->nblocks * 64;
    const bool first = p.cg->iter == 0;
    const float beta = (float)p.cg->beta;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        if (!(p.g_flags[n] & MPM_FLAG_INTERIOR_DEV)) continue;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const float r = p.g_r[a * p.gcap + n];
            p.g_s[a * p.gcap + n] = first ? r : beta * p.g_s[a * p.gcap + n] + r;
        }
    }
}

// q = dt^2/m q + s, Project(q), s.q
template <int D>
__global__ __launch_bounds__(WG) void k_cg_node1(KP p) {
    if (p.ctr->overflow || p.cg->done) return;
    const int nn = p.ctr->nblocks * 64;
    const float sds = p.dt * p.dt;
    double sum = 0.0, mx = 0.0;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float q[D];
#pragma unroll
        for (int a = 0; a < D; ++a) q[a] = p.g_q[a * p.gcap + n];
        if (interior) {
            const float c = sds / p.g_m[n];
#pragma unroll
            for (int a = 0; a < D; ++a) q[a] = c * q[a] + p.g_s[a * p.gcap + n];
        }
        if (project_zero<D>(p, n)) {
#pragma unroll
            for (int a = 0; a < D; ++a) q[a] = 0.0f;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) p.g_q[a * p.gcap + n] = q[a];
        if (interior) {
            const double m = (double)p.g_m[n];
            double sq = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) sq += (double)p.g_s[a * p.gcap + n] * ((double)q[a] * m);
            sum += sq;
        }
    }
    wg_reduce2(sum, mx);
    if (publish_and_check_last(p, sum, mx)) {
        gather_partials(p, sum, mx);
        if (threadIdx.x == 0) {
            CGState* cg = p.cg;
            cg->multiplies += 1;
            cg->sq = sum;
            cg->alpha = sum != 0.0 ? cg->rho / sum : (double)FLT_MAX;
            *p.sync = 0u;
        }
    }
}

// x += alpha s, r -= alpha q, then the convergence test of the next iteration
template <int D>
__global__ __launch_bounds__(WG) void k_cg_node2(KP p) {
    if (p.ctr->overflow || p.cg->done) return;
    const int nn = p.ctr->nblocks * 64;
    const float alpha = (float)p.cg->alpha;
    double sum = 0.0, mx = 0.0;
    for (int n = blockIdx.x * WG + threadIdx.x; n < nn; n += gridDim.x * WG) {
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        if (interior) {
            const double m = (double)p.g_m[n];
            double rr = 0.0, r2 = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float s = p.g_s[a * p.gcap + n], q = p.g_q[a * p.gcap + n];
                p.g_vs[a * p.gcap + n] = alpha * s + p.g_vs[a * p.gcap + n];
                const float r = -alpha * q + p.g_r[a * p.gcap + n];
                p.g_r[a * p.gcap + n] = r;
                rr += (double)r * ((double)r * m);
                r2 += (double)r * (double)r;
            }
            sum += rr;
            mx = fmax(mx, r2);
        }
#pragma unroll
        for (int a = 0; a < D; ++a) p.g_q[a * p.gcap + n] = 0.0f;
    }
    wg_reduce2(sum, mx);
    if (publish_and_check_last(p, sum, mx)) {
        gather_partials(p, sum, mx);
        if (threadIdx.x == 0) {
            CGState* cg = p.cg;
            cg->iter += 1;
            cg->conv = sqrt(mx);
            cg->rr = sum;
            if (cg->conv <= p.cg_tol && (cg->iter >= p.cg_min || cg->conv < (double)FLT_EPSILON)) {
                cg->done = 1;
                cg->converged = 1;
            } else if (cg->iter == p.cg_max) {
                cg->done = 1;
            } else {
                cg->rho_old = cg->rho;
                cg->rho = sum;
                cg->beta = cg->rho / cg->rho_old;
            }
            *p.sync = 0u;
        }
    }
}

// ============================================================================
// G2P (MPM_Example.cpp:631-660): PIC/FLIP velocity, F and X update, removal
// ============================================================================
template <int D>
__global__ __launch_bounds__(WG) void k_g2p(KP p) {
    using Dm = Dims<D>;
    constexpr int TN = Dm::TN, NBR = Dm::NBR, DD = D * D;
    __shared__ float vt[2 * D][TN];
    __shared__ int nbr[NBR];
    if (p.ctr->overflow) return;
    const int nblocks = p.ctr->nblocks;
    const float dt = p.dt, flip = p.flip;
    const float lo = 0.0f + 0.5f * p.dx, hi = 1.0f - 0.5f * p.dx;
    for (int b = xcd_wg(); b < nblocks; b += gridDim.x) {
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        slot_coords<D>(p.block_slot[b], p.nb, bb);
        __syncthreads();
        const float* src[2 * D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            src[a] = p.g_vs + a * p.gcap;
            src[D + a] = p.g_v + a * p.gcap;
        }
        gather_tile<D, 2 * D>(p, nbr, vt, src);
        __syncthreads();
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + threadIdx.x; i < i1; i += WG) {
            float w[D][3], dw[D][3], X[D];
            int l[D];
            if (!particle_weights<D>(p, i, bb, w, dw, l, X)) continue;
            float vpic[D], vflip[D], G[DD];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                vpic[a] = 0.0f;
                vflip[a] = p.pcur[(PF<D>::V + a) * p.pcap + i];
            }
#pragma unroll
            for (int k = 0; k < DD; ++k) G[k] = 0.0f;
#pragma unroll
            for (int s = 0; s < Dm::ST; ++s) {
                int off[D];
                const int t = tile_index<D>(l, s, off);
                const float wt = stencil_weight<D>(w, off);
                if (wt > 0.0f) {
                    float g[D];
                    stencil_grad<D>(w, dw, off, g);
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        const float vs = vt[a][t], v = vt[D + a][t];
                        vpic[a] += wt * vs;
                        vflip[a] += wt * (vs - v);
#pragma unroll
                        for (int c = 0; c < D; ++c) G[a * D + c] += vs * g[c];
                    }
                }
            }
            float F[DD];
#pragma unroll
            for (int k = 0; k < DD; ++k) F[k] = p.pcur[(PF<D>::F + k) * p.pcap + i];
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    float s = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) s += (dt * G[r * D + k]) * F[k * D + c];
                    p.pcur[(PF<D>::F + r * D + c) * p.pcap + i] = F[r * D + c] + s;
                }
            bool inside = true;
            float v2 = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float V = vflip[a] * flip + vpic[a] * (1.0f - flip);
                const float Xn = X[a] + vpic[a] * dt;
                p.pcur[(PF<D>::V + a) * p.pcap + i] = V;
                p.pcur[(PF<D>::X + a) * p.pcap + i] = Xn;
                inside = inside && Xn >= lo && Xn <= hi;
                v2 += V * V;
            }
            if (!inside) {
                p.keycur[i] = -1;
                atomicAdd(&p.ctr->n_removed, 1);
            } else atomicMax(&p.ctr->max_v2, __float_as_uint(v2));
        }
        __syncthreads();
    }
}

template <int D>
__global__ __launch_bounds__(WG) void k_max_speed(KP p) {
    const int n = p.ctr->n_live;
    unsigned best = 0u;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        if (p.keycur[i] < 0) continue;
        float v2 = 0.0f;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const float v = p.pcur[(PF<D>::V + a) * p.pcap + i];
            v2 += v * v;
        }
        best = max(best, __float_as_uint(v2));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, off));
    if ((threadIdx.x & 63) == 0 && best) atomicMax(&p.ctr->max_v2, best);
}

// ============================================================================
// launchers
// ============================================================================
template <int D> void launch_bin(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_bin<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_scan(const KP& p, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(p.nparts), dim3(WG), 0, s, p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(WG), 0, s, p);
    hipLaunchKernelGGL(k_scan_apply, dim3(p.nparts), dim3(WG), 0, s, p);
}
template <int D> void launch_block_setup(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_block_setup<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_reorder(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_reorder<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_p2g(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_p2g<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_normalize(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_normalize<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_init_volume(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_init_volume<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_particle_force(const KP& p, hipStream_t s, int grid, bool constitutive, bool force) {
    if (constitutive && force) hipLaunchKernelGGL((k_particle_force<D, true, true>), dim3(grid), dim3(WG), 0, s, p);
    else if (constitutive) hipLaunchKernelGGL((k_particle_force<D, true, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (force) hipLaunchKernelGGL((k_particle_force<D, false, true>), dim3(grid), dim3(WG), 0, s, p);
}
template <int D> void launch_grid_update(const KP& p, hipStream_t s, int grid, bool force, bool collision, bool cg_init) {
    if (force && collision && cg_init) hipLaunchKernelGGL((k_grid_update<D, true, true, true>), dim3(grid), dim3(WG), 0, s, p);
    else if (force && !collision && !cg_init) hipLaunchKernelGGL((k_grid_update<D, true, false, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (!force && collision && !cg_init) hipLaunchKernelGGL((k_grid_update<D, false, true, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (!force && !collision && cg_init) hipLaunchKernelGGL((k_grid_update<D, false, false, true>), dim3(grid), dim3(WG), 0, s, p);
    else if (force && collision && !cg_init) hipLaunchKernelGGL((k_grid_update<D, true, true, false>), dim3(grid), dim3(WG), 0, s, p);
}
template <int D> void launch_cg_force(const KP& p, hipStream_t s, int grid, bool restart) {
    if (restart) hipLaunchKernelGGL((k_cg_force<D, true>), dim3(grid), dim3(WG), 0, s, p);
    else hipLaunchKernelGGL((k_cg_force<D, false>), dim3(grid), dim3(WG), 0, s, p);
}
template <int D> void launch_cg_restart_node(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_cg_restart_node<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_cg_update_s(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_cg_update_s<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_cg_node1(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_cg_node1<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_cg_node2(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_cg_node2<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_g2p(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_g2p<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_max_speed(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_max_speed<D>, dim3(grid), dim3(WG), 0, s, p); }

#define MPM_INSTANTIATE(D)                                                                   \
    template void launch_bin<D>(const KP&, hipStream_t, int);                                \
    template void launch_scan<D>(const KP&, hipStream_t);                                    \
    template void launch_block_setup<D>(const KP&, hipStream_t, int);                        \
    template void launch_reorder<D>(const KP&, hipStream_t, int);                            \
    template void launch_p2g<D>(const KP&, hipStream_t, int);                                \
    template void launch_normalize<D>(const KP&, hipStream_t, int);                          \
    template void launch_init_volume<D>(const KP&, hipStream_t, int);                        \
    template void launch_particle_force<D>(const KP&, hipStream_t, int, bool, bool);         \
    template void launch_grid_update<D>(const KP&, hipStream_t, int, bool, bool, bool);      \
    template void launch_cg_force<D>(const KP&, hipStream_t, int, bool);                     \
    template void launch_cg_restart_node<D>(const KP&, hipStream_t, int);                    \
    template void launch_cg_update_s<D>(const KP&, hipStream_t, int);                        \
    template void launch_cg_node1<D>(const KP&, hipStream_t, int);                           \
    template void launch_cg_node2<D>(const KP&, hipStream_t, int);                           \
    template void launch_g2p<D>(const KP&, hipStream_t, int);                                \
    template void launch_max_speed<D>(const KP&, hipStream_t, int);
MPM_INSTANTIATE(2)
MPM_INSTANTIATE(3)

}  // namespace mpm
