// gfx950 kernels of the MPM snow substep, for a batch of scenes.
//
// Layout in HBM:
//   particles  field-major FP32 arrays (X[D], V[D], F[D*D], Jp, m, V0) of
//              capacity pcap, double buffered, with a scene id per particle;
//              after binning they are grouped by home block (so by scene,
//              then block, then cell) and every block's particles are one
//              contiguous range.
//   grid       block-sparse: each scene's padded domain is cut into 4^3
//              (3D) or 8^2 (2D) node blocks; the scenes' dense slot tables
//              are concatenated and map block coordinates to a compacted
//              ordinal; every node channel is an array of ordinal*64 + local
//              node entries.  Scene parameters are looked up per block.
//   tiles      per active block, the (B+2)^D node tile (block plus a
//              one-node halo) of each scattered channel.
// Block kernels (P2G, force, CG operator, G2P) run one work-group per
// active block.  A block's particles touch only its nodes and the one-node
// halo, so scatters accumulate into LDS tiles and the work-group writes
// the whole tile to the tile buffer with plain stores; the next node kernel
// sums, for every node, the (at most 2^D) tiles that cover it in a fixed
// order.  There are no global atomics on the scatter path, and the
// cross-block sum is deterministic.  Gathers load the tile once from the
// 3^D neighbour blocks.  Node kernels run one wave per block; reductions
// (CG dot products and norms, G2P's max speed) write one partial per block,
// and a one-work-group-per-scene kernel sums a scene's block partials in a
// fixed order.
#include "mpm_kernels.h"

namespace mpm {

template <int D>
struct PF {
    // MAT: the particle's material entry (int bits); J holds the density of EOS particles
    static constexpr int X = 0, V = D, F = 2 * D, J = 2 * D + D * D, M = J + 1, VOL = J + 2, MAT = J + 3, N = J + 4;
};

// Work-groups are dispatched round-robin over the 8 XCDs (work-group w runs
// on XCD w % 8).  Block loops give each XCD one contiguous eighth of the
// active blocks, so every XCD has work whatever the block count, and
// neighbouring blocks (which share halo nodes) meet in the same L2.
// Requires gridDim.x % 8 == 0.
struct XcdRange {
    int first, end, step;
};
__device__ inline XcdRange xcd_blocks(int n) {
    const int x = (int)(blockIdx.x & 7), k = (int)(blockIdx.x >> 3);
    const int lo = (int)(((int64_t)n * x) >> 3), hi = (int)(((int64_t)n * (x + 1)) >> 3);
    return {lo + k, hi, (int)(gridDim.x >> 3)};
}
// Node loops: one wave per block (lane = node), over the same per-XCD block
// range as the block kernels, so a node kernel reads the tiles its XCD's
// block kernel wrote.
__device__ inline XcdRange xcd_wave_blocks(int n) {
    constexpr int W = WG / 64;
    const int x = (int)(blockIdx.x & 7), k = (int)(blockIdx.x >> 3);
    const int lo = (int)(((int64_t)n * x) >> 3), hi = (int)(((int64_t)n * (x + 1)) >> 3);
    return {lo + k * W + (int)(threadIdx.x >> 6), hi, (int)(gridDim.x >> 3) * W};
}

// Active block b holds no particles (a halo block of its neighbours).
__device__ inline bool block_empty(const KP& p, int b) { return p.block_pstart[b + 1] == p.block_pstart[b]; }

// Scene of a global slot (slot ranges of an arena are not in scene order).
__device__ inline int scene_of_slot(const KP& p, int slot) { return p.slot_scene[slot]; }

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

// Coordinates of active block b in its scene's slot grid (written by
// k_block_setup; a load instead of the slot index's divisions).
template <int D>
__device__ inline void block_coords(const KP& p, int b, int* bb) {
#pragma unroll
    for (int a = 0; a < D; ++a) bb[a] = p.block_org[b * 3 + a];
}

template <int D>
__device__ inline void local_coords(int l, int* lc) {
    if constexpr (D == 3) { lc[0] = l >> 4; lc[1] = (l >> 2) & 3; lc[2] = l & 3; }
    else { lc[0] = l >> 3; lc[1] = l & 7; }
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

// stencil offsets of node s (last axis fastest) -> LDS tile index of that node for local cell l
template <int D>
__device__ inline int tile_index(const int* l, int s, int* off) {
    using Dm = Dims<D>;
    if constexpr (D == 3) {
        off[0] = s / 9;
        off[1] = (s / 3) % 3;
        off[2] = s % 3;
        return (l[0] + off[0]) * Dm::LX + (l[1] + off[1]) * Dm::LY + (l[2] + off[2]);
    } else {
        off[0] = s / 3;
        off[1] = s % 3;
        return (l[0] + off[0]) * Dm::LX + (l[1] + off[1]);
    }
}

// LDS index of tile node tc (tile coordinates 0..TE-1)
template <int D>
__device__ inline int lds_index(const int* tc) {
    using Dm = Dims<D>;
    if constexpr (D == 3) return tc[0] * Dm::LX + tc[1] * Dm::LY + tc[2];
    else return tc[0] * Dm::LX + tc[1];
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
__device__ inline bool particle_weights(const KP& p, const SceneParams& sc, int i, const int* bb, float (*w)[3],
                                        float (*dw)[3], int* l, float* X) {
    constexpr int B = Dims<D>::B;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        X[a] = p.pcur[(PF<D>::X + a) * p.pcap + i];
        int c;
        bspline_weights(X[a], sc.dx, sc.inv_dx, c, w[a], dw[a]);
        l[a] = c + B - B * bb[a];
        ok = ok && l[a] >= 0 && l[a] < B;
    }
    return ok;
}

// Sum, for node n, of channel c over every tile covering it, in a fixed
// order.  Blocks without particles write no tile (their contribution is
// zero) and are absent from the tile neighbour table.  With crop_ghost_x, nodes whose x index is outside [0, counts-1]
// receive nothing (the reference's slab crop,
// Tools/Cropped_Influence_Iterator.h:31-42).
template <int D, int C>
__device__ inline void gather_node_tiles(const KP& p, const SceneParams& sc, int n, float* out) {
    using Dm = Dims<D>;
    constexpr int B = Dm::B, TE = Dm::TE, TN = Dm::TN, NBR = Dm::NBR;
    const int b = n >> 6;
    int lc[D];
    local_coords<D>(n & 63, lc);
    // per axis: option 0 is the node's own block; a node on a block face
    // also lies in the face neighbour's tile (option 1)
    int od1[D], td0[D], td1[D], face = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        td0[a] = lc[a] + 1;
        face |= (lc[a] == 0 || lc[a] == B - 1) ? 1 << (D - 1 - a) : 0;
        od1[a] = lc[a] == 0 ? 0 : 2;
        td1[a] = lc[a] == 0 ? TE - 1 : 0;
    }
    // every candidate tile's neighbour entry, then every tile value, is
    // loaded before the first sum (no load waits on another candidate's);
    // candidate m takes option 1 on the axes of its set bits, and the sum
    // runs over the candidates in lexicographic option order
    constexpr int NC = 1 << D;
    const int* nbr = p.block_tnbr + (int64_t)b * NBR;
    int blk[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        int o = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) o = o * 3 + (((m >> (D - 1 - a)) & 1) ? od1[a] : 1);
        const int e = nbr[o];  // o is an entry of the table whether or not it is used
        blk[m] = (m & ~face) == 0 ? e : -1;
    }
    // a candidate on the node's faces loads even where its block wrote no
    // tile (from block 0's slot, in bounds, and not summed), so that no sum
    // shares its load's condition and none is scheduled ahead of the loads
    float v[NC][C];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        int t = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) t = t * TE + (((m >> (D - 1 - a)) & 1) ? td1[a] : td0[a]);
#pragma unroll
        for (int c = 0; c < C; ++c) v[m][c] = 0.0f;
        if ((m & ~face) == 0) {
            const float* src = p.tiles + (int64_t)max(blk[m], 0) * C * TN + t;
#pragma unroll
            for (int c = 0; c < C; ++c) v[m][c] = src[c * TN];
        }
    }
    // the crop test's loads go out with the tile loads (a cropped node sums
    // nothing)
    bool keep = true;
    {
        const int crop = sc.crop_ghost_x, counts = sc.counts;
        int bb[D];
        block_coords<D>(p, b, bb);
        const int ix = B * bb[0] + lc[0] - B;
        keep = (crop == 0) | ((ix >= 0) & (ix <= counts - 1));
    }
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = 0.0f;
#pragma unroll
    for (int m = 0; m < NC; ++m)
        if (keep && blk[m] >= 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) out[c] += v[m][c];
        }
}

__device__ inline bool inside_plane(const SceneParams& sc, int id, const float* x, int D) {
    float s = 0.0f;
    for (int a = 0; a < D; ++a) s += (x[a] - sc.ps[id][a]) * sc.pn[id][a];
    return s < 0.0f;
}

template <int D>
__device__ inline void node_center(const KP& p, const SceneParams& sc, int n, float* x) {
    constexpr int B = Dims<D>::B;
    int bb[D], lc[D];
    block_coords<D>(p, n >> 6, bb);
    local_coords<D>(n & 63, lc);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const int i = B * bb[a] + lc[a] - B;
        x[a] = ((float)i + 0.5f) * sc.dx;
    }
}

// ============================================================================
// Binning: key = global slot*64 + cell of the home cell, rank within the
// cell, activation marks of the home block's 3^D neighbourhood.  Empty and
// removed entries carry key -1.
// ============================================================================
template <int D>
__global__ __launch_bounds__(WG) void k_bin(KP p) {
    constexpr int B = Dims<D>::B;
    const int n = p.ctr->n_live;
    const int lane = threadIdx.x & 63;
    // whole waves iterate together so the ballots below are convergent
    for (int base = blockIdx.x * WG + (threadIdx.x & ~63); base < n; base += gridDim.x * WG) {
        const int i = base + lane;
        // the key, scene and position load together (a lane past the end
        // reads the last particle and is not valid), then the scene's
        // parameters (scene 0's for a lane that is not live); no branch
        // separates them, so no load waits on another
        const int ic = min(i, n - 1);
        const int key0 = p.keycur[ic], si0 = p.scur[ic];
        float X[D];
#pragma unroll
        for (int a = 0; a < D; ++a) X[a] = p.pcur[(PF<D>::X + a) * p.pcap + ic];
        const bool live = i < n && key0 >= 0;
        const int si = live ? si0 : 0;
        const SceneParams& sc = p.scn[si];
        const float inv_dx = sc.inv_dx;
        const int counts = sc.counts, nb = sc.nb, slot_base = sc.slot_base;
        bool finite = true;
#pragma unroll
        for (int a = 0; a < D; ++a) finite = finite & (bool)isfinite(X[a]);
        int c[D];
        bool valid = live && finite;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            c[a] = finite ? cell_of(X[a], inv_dx) : -1;
            valid = valid & (c[a] >= 0) & (c[a] <= counts - 1);
        }
        if (live && !valid) {
            p.keycur[i] = -1;
            if (!finite) atomicOr(&p.sst[si].nonfinite, 1);
        }
        int slot = -1, key = -1;
        if (valid) {
            int local = 0, cell = 0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                local = local * nb + (c[a] + B) / B;
                cell = cell * B + (c[a] + B) % B;
            }
            slot = slot_base + local;
            key = slot * 64 + cell;
        }
        // wave-aggregated atomics: one per distinct (block, cell) for the rank in
        // the cell, one per distinct block for the block total.  The groups are
        // found first (ballots only), then every group's first lane adds its
        // group's size at once, so a wave waits for one round of atomics.
        int leader = 0, in_group = 0, group_size = 0;
        unsigned long long todo = __ballot(valid);
        while (todo) {
            const int l = __ffsll((long long)todo) - 1;
            const int lkey = __shfl(key, l);
            const unsigned long long grp = __ballot(valid && key == lkey);
            if (valid && key == lkey) {
                leader = l;
                in_group = __popcll(grp & ((1ull << lane) - 1ull));
            }
            if (lane == l) group_size = __popcll(grp);
            todo &= ~grp;
        }
        int r0 = 0;
        if (valid && lane == leader) r0 = atomicAdd(&p.cell_count[key], group_size);
        const int rank = __shfl(r0, leader) + in_group;
        todo = __ballot(valid);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int lslot = __shfl(slot, leader);
            const unsigned long long grp = __ballot(valid && slot == lslot);
            if (lane == leader) atomicAdd(&p.slot_count[lslot], __popcll(grp));
            todo &= ~grp;
        }
        if (valid) {
            p.prank[i] = rank;
            p.keycur[i] = key;
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < D; ++a) { lo[a] = (c[a] + B - 1) / B; hi[a] = (c[a] + B + 1) / B; }
            const int sb = slot_base;
            if constexpr (D == 3) {
                for (int x = lo[0]; x <= hi[0]; ++x)
                    for (int y = lo[1]; y <= hi[1]; ++y)
                        for (int z = lo[2]; z <= hi[2]; ++z) p.slot_touched[sb + (x * nb + y) * nb + z] = 1;
            } else {
                for (int x = lo[0]; x <= hi[0]; ++x)
                    for (int y = lo[1]; y <= hi[1]; ++y) p.slot_touched[sb + x * nb + y] = 1;
            }
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

// Slots scanned: [0, slot_hw) -- an arena keeps its scenes' ranges low.
__device__ inline int scanned_slots(const KP& p) { return min(p.nslots, p.ctr->slot_hw); }

// A thread's SCAN_PER_THREAD consecutive slots from `base`: marks and particle
// counts, one int4 each where all of them are below ns (base is a multiple
// of 4 and the arrays are 256-byte aligned), zeros past ns.
__device__ inline void load_slots(const KP& p, int base, int ns, int* act, int* cnt) {
    static_assert(SCAN_PER_THREAD == 4, "one int4 per array");
    if (base + SCAN_PER_THREAD <= ns) {
        const int4 a = *reinterpret_cast<const int4*>(p.slot_touched + base);
        const int4 c = *reinterpret_cast<const int4*>(p.slot_count + base);
        act[0] = a.x; act[1] = a.y; act[2] = a.z; act[3] = a.w;
        cnt[0] = c.x; cnt[1] = c.y; cnt[2] = c.z; cnt[3] = c.w;
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k) {
            const int s = base + k;
            act[k] = s < ns ? p.slot_touched[s] : 0;
            cnt[k] = s < ns ? p.slot_count[s] : 0;
        }
    }
}

__global__ __launch_bounds__(WG) void k_scan_reduce(KP p) {
    const int ns = scanned_slots(p);
    if ((int64_t)blockIdx.x * SCAN_TILE >= ns) {  // beyond the scanned slots: empty part
        if (threadIdx.x == 0) {
            p.scan_part[blockIdx.x] = 0;
            p.scan_part[p.nparts + blockIdx.x] = 0;
        }
        return;
    }
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    int at[SCAN_PER_THREAD], ct[SCAN_PER_THREAD];
    load_slots(p, base, ns, at, ct);
    int act = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        act += at[k] != 0;
        cnt += ct[k];
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
        if (ta > p.bcap) {
            // the substep becomes a no-op: particles keep their order and state
            p.ctr->overflow = 1;
            p.ctr->nblocks = 0;
        } else {
            p.ctr->overflow = 0;  // a binning that fits clears an earlier overflow
            p.ctr->n_live = tc;
            p.ctr->nblocks = ta;
            p.block_pstart[ta] = tc;
        }
    }
}

__global__ __launch_bounds__(WG) void k_scan_apply(KP p) {
    const int ns = scanned_slots(p);
    if ((int64_t)blockIdx.x * SCAN_TILE >= ns) return;
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    int act[SCAN_PER_THREAD], cnt[SCAN_PER_THREAD];
    load_slots(p, base, ns, act, cnt);
    int sa = 0, sc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        act[k] = act[k] != 0;
        sa += act[k];
        sc += cnt[k];
    }
    int ta, tc;
    int ra = wg_exclusive_scan(sa, ta) + p.scan_part[blockIdx.x];
    int rc = wg_exclusive_scan(sc, tc) + p.scan_part[p.nparts + blockIdx.x];
    // after an overflow no block is set up, so the binning's cell counters of
    // the touched slots are cleared here instead of by the block set-up
    const bool overflow = p.ctr->overflow != 0;
    int sact[SCAN_PER_THREAD];
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        const int s = base + k;
        sact[k] = act[k] ? ra : -1;
        if (act[k] && s < ns) {
            if (ra < p.bcap) {
                p.block_slot[ra] = s;
                p.block_pstart[ra] = rc;
            }
            if (overflow)
                for (int cell = 0; cell < 64; ++cell) p.cell_count[s * 64 + cell] = 0;
        }
        ra += act[k];
        rc += cnt[k];
    }
    // the thread's slots: active index, and the binning's marks cleared
    if (base + SCAN_PER_THREAD <= ns) {
        *reinterpret_cast<int4*>(p.slot_active + base) = make_int4(sact[0], sact[1], sact[2], sact[3]);
        *reinterpret_cast<int4*>(p.slot_touched + base) = make_int4(0, 0, 0, 0);
        *reinterpret_cast<int4*>(p.slot_count + base) = make_int4(0, 0, 0, 0);
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD; ++k) {
            const int s = base + k;
            if (s < ns) {
                p.slot_active[s] = sact[k];
                p.slot_touched[s] = 0;
                p.slot_count[s] = 0;
            }
        }
    }
}

// Per active block (one wave, lane = cell): its scene, the neighbour table
// (inside the scene's slot range), and the cell ranges -- an exclusive scan
// of the block's 64 cell counts -- after which the cell counters are
// cleared for the next binning.
template <int D>
__global__ __launch_bounds__(WG) void k_block_setup(KP p) {
    using Dm = Dims<D>;
    const int nblocks = p.ctr->nblocks;
    const int group = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = blockIdx.x * (WG / 64) + group; b < nblocks; b += gridDim.x * (WG / 64)) {
        const int slot = p.block_slot[b];
        const int si = scene_of_slot(p, slot);
        const SceneParams& sc = p.scn[si];
        if (lane == 0) p.block_scene[b] = si;
        if (lane < Dm::NBR) {
            int bb[D];
            slot_coords<D>(slot - sc.slot_base, sc.nb, bb);
            if (lane < D) p.block_org[b * 3 + lane] = bb[lane];
            int o[3];
            if constexpr (D == 3) { o[0] = lane / 9; o[1] = (lane / 3) % 3; o[2] = lane % 3; }
            else { o[0] = lane / 3; o[1] = lane % 3; }
            int s = 0;
            bool in = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const int c = bb[a] + o[a] - 1;
                in = in && c >= 0 && c < sc.nb;
                s = s * sc.nb + c;
            }
            const int act = in ? p.slot_active[sc.slot_base + s] : -1;
            p.block_nbr[b * Dm::NBR + lane] = act;
            // tiles exist only for blocks that hold particles
            p.block_tnbr[b * Dm::NBR + lane] = (act >= 0 && p.block_pstart[act + 1] > p.block_pstart[act]) ? act : -1;
        }
        const int cnt = p.cell_count[slot * 64 + lane];
        int x = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        const int start = p.block_pstart[b] + x - cnt;
        p.cell_start[b * 65 + lane] = start;
        if (lane == 63) p.cell_start[b * 65 + 64] = start + cnt;
        int full = cnt;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) full = max(full, __shfl_xor(full, off));
        if (lane == 0) p.block_full[b] = full;
        p.cell_count[slot * 64 + lane] = 0;
    }
}

// Per scene: its active block range and particle range (active blocks are
// in slot order, so each scene's blocks are contiguous).
__device__ inline void scene_ranges(const KP& p, int bid) {
    const int si = bid * WG + (int)threadIdx.x;
    if (si >= p.nscenes) return;
    const int nb = p.ctr->nblocks;
    auto lower = [&](int slot) {
        int lo = 0, hi = nb;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (p.block_slot[mid] < slot) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const SceneParams& sc = p.scn[si];
    const int b0 = lower(sc.slot_base), b1 = lower(sc.slot_base + sc.nslots);
    SceneState& st = p.sst[si];
    st.block_begin = b0;
    st.block_end = b1;
    st.pbegin = nb > 0 ? p.block_pstart[b0] : 0;
    st.pend = nb > 0 ? p.block_pstart[b1] : 0;
}

// One work-group per XCD range of the active blocks (the ranges of
// xcd_blocks) and per key: the range's blocks by descending particle count
// (bucketed by eight; work-groups 0-7) and by descending count of their
// fullest cell (work-groups 8-15), each a counting sort whose order within
// a bucket is that of the LDS atomics.  The CG operator's time per block
// follows its fullest cell (its lane-per-cell passes are as many as that
// cell's particles); P2G, the constitutive/force kernel and G2P measured
// fastest in particle-count order.  Block loops that follow the hardware's
// dispatch order then start the long blocks first and leave short ones for
// the tail.  Only which work-group computes a block changes, never a
// block's result.
__device__ inline void block_order(const KP& p, int bid) {
    constexpr int NBK = WG;  // buckets, one per thread for the scan
    __shared__ int hist[NBK];
    const int n = p.ctr->nblocks, x = bid & 7, by_cell = bid >> 3;
    const int lo = (int)(((int64_t)n * x) >> 3), hi = (int)(((int64_t)n * (x + 1)) >> 3);
    int* order = p.block_order + (by_cell ? p.bcap : 0);
    auto key = [&](int b) {  // the longest first
        const int w = by_cell ? p.block_full[b] : (p.block_pstart[b + 1] - p.block_pstart[b]) >> 3;
        return NBK - 1 - min(NBK - 1, w);
    };
    hist[threadIdx.x] = 0;
    __syncthreads();
    for (int b = lo + (int)threadIdx.x; b < hi; b += WG) atomicAdd(&hist[key(b)], 1);
    __syncthreads();
    int total;
    const int off = wg_exclusive_scan(hist[threadIdx.x], total);
    __syncthreads();
    hist[threadIdx.x] = off;
    __syncthreads();
    for (int b = lo + (int)threadIdx.x; b < hi; b += WG) order[lo + atomicAdd(&hist[key(b)], 1)] = b;
}

// scatter into (block, cell) order
// The binning ranks come from atomics, so the order they give within a cell
// depends on arrival.  Each particle writes its index into its cell's range
// at that rank; k_reorder then places it by the number of indices in the
// range below its own, which orders every cell by the particles' previous
// positions (a stable sort by cell) and makes the layout, and every sum over
// a cell's particles, the same from run to run.
__device__ inline void cell_order(const KP& p, int bid, int nbid) {
    if (p.ctr->overflow != 0) return;
    const int n = p.ctr->n_prev;
    for (int i = bid * WG + (int)threadIdx.x; i < n; i += nbid * WG) {
        const int key = p.keycur[i];
        if (key < 0) continue;
        p.cell_order[p.cell_start[p.slot_active[key >> 6] * 65 + (key & 63)] + p.prank[i]] = i;
    }
}

// The three set-up steps that follow the block set-up and do not depend on
// one another, in one launch: work-groups 0-15 sort the blocks (the XCD of
// work-group w is w mod 8, as for the block kernels), the next ceil(scenes /
// WG) find the scenes' ranges, the rest scatter the particles' indices into
// their cells.
constexpr int ORDER_GROUPS = 16;
__global__ __launch_bounds__(WG) void k_setup_order(KP p) {
    const int sr = (p.nscenes + WG - 1) / WG;
    const int bid = (int)blockIdx.x;
    if (bid < ORDER_GROUPS) block_order(p, bid);
    else if (bid < ORDER_GROUPS + sr) scene_ranges(p, bid - ORDER_GROUPS);
    else cell_order(p, bid - ORDER_GROUPS - sr, (int)gridDim.x - ORDER_GROUPS - sr);
}

template <int D>
__global__ __launch_bounds__(WG) void k_reorder(KP p) {
    const int n = p.ctr->n_prev;
    const bool overflow = p.ctr->overflow != 0;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        // the particle's state loads with its key, before the destination is
        // known (the buffers may alias as far as the compiler knows, so a load
        // after a store would wait for it).  Only a placed particle (a live key
        // and no overflow) reads the cell tables; a removed one (key < 0) or an
        // identity copy reads cell 0's bounds (the table always has them: slot
        // tables hold -1, and after an overflow indices past the capacity, so
        // they are not read for these) and a removed one stores nothing.
        const int key = p.keycur[i];
        float v[PF<D>::N];
#pragma unroll
        for (int f = 0; f < PF<D>::N; ++f) v[f] = p.pcur[f * p.pcap + i];
        const int id = p.idcur[i], si = p.scur[i];
        const bool live = overflow || key >= 0;
        const bool placed = !overflow && key >= 0;
        const int cs = placed ? p.slot_active[key >> 6] * 65 + (key & 63) : 0;
        const int start = p.cell_start[cs], end = p.cell_start[cs + 1];
        // rank: indices of the cell below i, eight loads in flight at a time
        // (an index past the range reads its last entry and is not counted)
        constexpr int U = 8;
        int rank = 0;
        if (placed) {
            for (int j0 = start; j0 < end; j0 += U) {
                int e[U];
#pragma unroll
                for (int u = 0; u < U; ++u) e[u] = p.cell_order[min(j0 + u, end - 1)];
#pragma unroll
                for (int u = 0; u < U; ++u) rank += (j0 + u < end && e[u] < i) ? 1 : 0;
            }
        }
        // overflow: identity copy, the buffer swap leaves the state unchanged
        const int dest = overflow ? i : start + rank;
        if (!live) continue;
#pragma unroll
        for (int f = 0; f < PF<D>::N; ++f) p.palt[f * p.pcap + dest] = v[f];
        p.idalt[dest] = id;
        p.salt[dest] = si;
        p.keyalt[dest] = overflow ? key : 0;
    }
}

// Block kernels run NW waves per active block.  In scatter kernels lane c of
// every wave owns cell c of the block (particles are in (block, cell) order,
// cell ranges in cell_start) and wave w takes the cell's particles w, w+NW,
// ...  All particles of a lane share one 3^D stencil and, for a fixed stencil
// offset, the 64 lanes of a wave address 64 distinct tile nodes.  Each wave
// therefore scatters into its own LDS tile with plain read-add-write steps
// (one stencil node per step; LDS operations of a wave execute in order), and
// the NW wave tiles are summed when the block's tile is stored.  No LDS or
// global atomics are used.
constexpr int NW = 4;
constexpr int BWG = 64 * NW;

// Compiler-only memory barrier between two read-add-write steps of a wave
// scatter: a lane's step s+1 may read the node another lane wrote in step s,
// so no LDS access may move across it (a wave's LDS operations execute in
// issue order).
#define MPM_STEP_FENCE() __asm__ __volatile__("" ::: "memory")

// The tile nodes a thread gathers and stores in a D-channel block kernel
// depend on its thread index only, so the kernel forms them once:
//   g: gather of tile node threadIdx.x: neighbour | local node << 5 | LDS
//      index << 11 (~0u for threads past the tile);
//   s: store of tile values t = threadIdx.x + k BWG (k < 3): the
//      channel-major LDS index of each, 10 bits apiece.
template <int D>
struct TileMaps {
    unsigned g, s;
    __device__ void init() {
        constexpr int TN = Dims<D>::TN, LTN = Dims<D>::LTN;
        static_assert(D * TN <= 3 * BWG && D * LTN <= 1024, "three 10-bit store slots");
        g = ~0u;
        if ((int)threadIdx.x < TN) {
            int nb, local, tc[D];
            tile_map<D>(threadIdx.x, nb, local, tc);
            g = (unsigned)nb | ((unsigned)local << 5) | ((unsigned)lds_index<D>(tc) << 11);
        }
        s = 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int t = threadIdx.x + k * BWG;
            if (t < D * TN) {
                const int c = t / TN;
                int nb, local, tc[D];
                tile_map<D>(t - c * TN, nb, local, tc);
                s |= (unsigned)(c * LTN + lds_index<D>(tc)) << (10 * k);
            }
        }
    }
};

// store_wave_tiles<D, D> with the thread's maps
template <int D>
__device__ inline void store_wave_tiles_mapped(const KP& p, int b, const float* lds, unsigned smap) {
    constexpr int TN = Dims<D>::TN, LTN = Dims<D>::LTN;
    float* dst = p.tiles + (int64_t)b * D * TN;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int t = threadIdx.x + k * BWG;
        if (t < D * TN) {
            const int li = (int)((smap >> (10 * k)) & 1023u);
            float v = 0.0f;
#pragma unroll
            for (int w = 0; w < NW; ++w) v += lds[w * D * LTN + li];
            dst[t] = v;
        }
    }
}

// Sum of the NW wave tiles (C channels each, LDS layout) of block b into
// its tile slot (compact layout, node (x,y,z) at (x*TE + y)*TE + z).
template <int D, int C>
__device__ inline void store_wave_tiles(const KP& p, int b, const float* lds) {
    constexpr int TN = Dims<D>::TN, LTN = Dims<D>::LTN;
    float* dst = p.tiles + (int64_t)b * C * TN;
    for (int t = threadIdx.x; t < C * TN; t += BWG) {
        const int c = t / TN;
        int nb, local, tc[D];
        tile_map<D>(t - c * TN, nb, local, tc);
        const int li = c * LTN + lds_index<D>(tc);
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; ++w) v += lds[w * C * LTN + li];
        dst[t] = v;
    }
}

// Weights of particle i; false if it is not in cell l of block bb (only after
// particles moved without re-binning: the particle is skipped).
template <int D>
__device__ inline bool cell_particle(const KP& p, const SceneParams& sc, int i, const int* bb, const int* l,
                                     float (*w)[3], float (*dw)[3], float* X) {
    int lp[D];
    if (!particle_weights<D>(p, sc, i, bb, w, dw, lp, X)) return false;
    bool same = true;
#pragma unroll
    for (int a = 0; a < D; ++a) same = same && lp[a] == l[a];
    return same;
}

// The record of an EOS particle: its CG force V0 K gamma rho^(gamma-1)
// (grad w . u) grad w (MPM_CG_System.h:169-174, 205-207) is A = c tr(T) I,
// taking the reference's matrix-plus-scalar as adding to the diagonal.  In
// the record's layout: U = I, every entry of the diagonal block c, no pair
// blocks (2D: c at (00,00), (00,11), (11,11)).
template <int D>
__device__ inline void store_operator_eos(const KP& p, int i, float c, const float* xo);

// ============================================================================
// Linearised force operator per particle.  MPM_CG_System::Force
// (MPM_CG_System.h:150-215) maps the node field u to forces through
//   T = sum_i grad w_ip (x) u_i,  dF = T^T F,  A = V0 F dP(dF)^T,
//   q_i += A^T grad w_ip.
// T -> A is linear and fixed for the whole solve, so it is evaluated once
// per substep as a D^2 x D^2 matrix (columns = A of the unit T's), which is
// symmetric (dP is the Hessian of the energy: V0 dP(dF):dF' = V0 dP(dF'):dF)
// and is stored after the particle's stencil offsets as its upper triangle.
// One CG application then costs D^4 multiply-adds per particle instead of
// re-deriving the stress state, and reads one contiguous record.
// ============================================================================
template <int D> struct OpRec;
// 3D: x[3] (the stencil offsets of the particle's position in its binned
// cell, in cells), q[4] (unit quaternion of U), the singular values S[3] and the
// four scalars A = 2 mu V0, B = lambda V0 (J - 1), P = B J, Q = lambda V0 J^2
// from which the 3x3 block on the diagonal entries and the three 2x2 pair
// blocks follow, pad 2 (64 B: four float4).
template <> struct OpRec<3> { static constexpr int W = 16; };  // 4 x float4
// 2D: x[2] (stencil offsets), the full symmetric 4x4 operator (upper triangle, 10).
template <> struct OpRec<2> { static constexpr int L = 10, W = 12; };  // 3 x float4
static_assert(OpRec<3>::W == op_record_floats(3) && OpRec<2>::W == op_record_floats(2), "operator record size");

template <int D>
__device__ constexpr int op_index(int r, int c) {
    // upper triangle, row-major, of a D^2 x D^2 matrix (r <= c)
    return r * D * D - (r * (r - 1)) / 2 + (c - r);
}

// A = V0 F dP(T^T F)^T, the reference's per-particle force map
template <int D>
__device__ inline void force_map(const StressState<D>& st, float V0, const float* T, float* A) {
    constexpr int DD = D * D;
    float dF[DD];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int cc = 0; cc < D; ++cc) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) s += T[k * D + r] * st.F[k * D + cc];
            dF[r * D + cc] = s;
        }
    float dP[DD];
    dP_of<D>(st, dF, dP);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int cc = 0; cc < D; ++cc) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) s += st.F[r * D + k] * dP[cc * D + k];
            A[r * D + cc] = s * V0;
        }
}

// rotation matrix (row-major) <-> unit quaternion (w, x, y, z)
__device__ inline void quat_of(const float* U, float* q) {
    const float tr = U[0] + U[4] + U[8];
    float w, x, y, z;
    if (tr > 0.0f) {
        const float s = 2.0f * sqrtf(tr + 1.0f);
        w = 0.25f * s; x = (U[7] - U[5]) / s; y = (U[2] - U[6]) / s; z = (U[3] - U[1]) / s;
    } else if (U[0] > U[4] && U[0] > U[8]) {
        const float s = 2.0f * sqrtf(1.0f + U[0] - U[4] - U[8]);
        w = (U[7] - U[5]) / s; x = 0.25f * s; y = (U[1] + U[3]) / s; z = (U[2] + U[6]) / s;
    } else if (U[4] > U[8]) {
        const float s = 2.0f * sqrtf(1.0f + U[4] - U[0] - U[8]);
        w = (U[2] - U[6]) / s; x = (U[1] + U[3]) / s; y = 0.25f * s; z = (U[5] + U[7]) / s;
    } else {
        const float s = 2.0f * sqrtf(1.0f + U[8] - U[0] - U[4]);
        w = (U[3] - U[1]) / s; x = (U[2] + U[6]) / s; y = (U[5] + U[7]) / s; z = 0.25f * s;
    }
    const float n = 1.0f / sqrtf(w * w + x * x + y * y + z * z);
    q[0] = w * n; q[1] = x * n; q[2] = y * n; q[3] = z * n;
}
__device__ inline void rot_of(const float* q, float* U) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    U[0] = 1.0f - 2.0f * (y * y + z * z); U[1] = 2.0f * (x * y - w * z);        U[2] = 2.0f * (x * z + w * y);
    U[3] = 2.0f * (x * y + w * z);        U[4] = 1.0f - 2.0f * (x * x + z * z); U[5] = 2.0f * (y * z - w * x);
    U[6] = 2.0f * (x * z - w * y);        U[7] = 2.0f * (y * z + w * x);        U[8] = 1.0f - 2.0f * (x * x + y * y);
}

// Ahat = U^T A U (3x3)
__device__ inline void rotate_in(const float* U, const float* A, float* out) {
    float t[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t[i * 3 + j] = A[i * 3 + 0] * U[0 * 3 + j] + A[i * 3 + 1] * U[1 * 3 + j] + A[i * 3 + 2] * U[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = U[0 * 3 + i] * t[0 * 3 + j] + U[1 * 3 + i] * t[1 * 3 + j] + U[2 * 3 + i] * t[2 * 3 + j];
}
// A = U Ahat U^T (3x3)
__device__ inline void rotate_out(const float* U, const float* Ah, float* out) {
    float t[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t[i * 3 + j] = U[i * 3 + 0] * Ah[0 * 3 + j] + U[i * 3 + 1] * Ah[1 * 3 + j] + U[i * 3 + 2] * Ah[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = t[i * 3 + 0] * U[j * 3 + 0] + t[i * 3 + 1] * U[j * 3 + 1] + t[i * 3 + 2] * U[j * 3 + 2];
}

// The record of particle i.  3D: in the frame of U (F = U S V^T, R = U V^T)
// the reference's force map T -> A = V0 F dP(T^T F)^T of fixed-corotated
// snow (MPM_CG_System.h:176-210, dP of MPM_Constitutive_Model.cpp) decouples:
// with T^ = U^T T U, dF^ = T^T S, the rotation term W^_ij = (dF^_ij -
// dF^_ji) / (s_i + s_j) and the derivative of cof(S), A^ = U^T A U = V0 S dP^T
// is a 3x3 block on the diagonal entries,
//   V0 (2 mu s_a^2 [c == a] + lambda J^2 + lambda J (J - 1) [c != a])
//   = diag(A s_a^2 - P) + (P + Q) 1 1^T,
// plus per off-diagonal pair (lo, hi), with s = s_lo + s_hi and k the third
// index, the symmetric block
//   [[A s_lo^2 (s - 1) / s,  s_lo s_hi (A / s - B s_k)],
//    [         ...,          A s_hi^2 (s - 1) / s   ]],
// where A = 2 mu V0, B = lambda V0 (J - 1), P = B J, Q = lambda V0 J^2.  The
// record holds S and these four scalars (formed in FP64, stored in FP32);
// apply_operator forms the block entries from them.
template <int D>
__device__ inline void store_operator(const KP& p, int i, const StressState<D>& st, float V0, const float* xo,
                                      const float* U, const double* S) {
    constexpr int DD = D * D;
    float rec[OpRec<D>::W];
#pragma unroll
    for (int k = 0; k < OpRec<D>::W; ++k) rec[k] = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a) rec[a] = xo[a];
    if constexpr (D == 3) {
        quat_of(U, rec + 3);
        const double v0 = (double)V0, mu = (double)st.mu, lam = (double)st.lambda;
        const double J = S[0] * S[1] * S[2];
        const double B = lam * v0 * (J - 1.0);
        rec[7] = (float)S[0];
        rec[8] = (float)S[1];
        rec[9] = (float)S[2];
        rec[10] = (float)(2.0 * mu * v0);
        rec[11] = (float)B;
        rec[12] = (float)(B * J);
        rec[13] = (float)(lam * v0 * J * J);
    } else {
        float L[DD][DD];
#pragma unroll
        for (int j = 0; j < DD; ++j) {
            float T[DD], A[DD];
#pragma unroll
            for (int k = 0; k < DD; ++k) T[k] = k == j ? 1.0f : 0.0f;
            force_map<D>(st, V0, T, A);
#pragma unroll
            for (int r = 0; r < DD; ++r) L[r][j] = A[r];
        }
#pragma unroll
        for (int r = 0; r < DD; ++r)
#pragma unroll
            for (int c = r; c < DD; ++c) rec[D + op_index<D>(r, c)] = 0.5f * (L[r][c] + L[c][r]);
    }
    float4* dst = reinterpret_cast<float4*>(p.pop + (int64_t)i * OpRec<D>::W);
#pragma unroll
    for (int k = 0; k < OpRec<D>::W / 4; ++k) dst[k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
}

// The 2x2 block of pair (lo, hi) with third index k applied to
// (Th_lo,hi, Th_hi,lo).
__device__ inline void apply_pair(float A, float B, float slo, float shi, float sk, float tlh, float thl, float& alh,
                                  float& ahl) {
    const float sm = slo + shi;
    const float inv = __builtin_amdgcn_rcpf(sm);
    const float g = A * (sm - 1.0f) * inv;
    const float a = g * slo * slo, c = g * shi * shi;
    const float b = slo * shi * (A * inv - B * sk);
    alh = a * tlh + b * thl;
    ahl = b * tlh + c * thl;
}

// A = L T from a particle's operator record
template <int D>
__device__ inline void apply_operator(const float* rec, const float* T, float* A) {
    constexpr int DD = D * D;
    if constexpr (D == 3) {
        const float4 r0 = *reinterpret_cast<const float4*>(rec);       // X0 X1 X2 q0
        const float4 r1 = *reinterpret_cast<const float4*>(rec + 4);   // q1 q2 q3 s0
        const float4 r2 = *reinterpret_cast<const float4*>(rec + 8);   // s1 s2 A B
        const float4 r3 = *reinterpret_cast<const float4*>(rec + 12);  // P Q - -
        const float q[4] = {r0.w, r1.x, r1.y, r1.z};
        const float s0 = r1.w, s1 = r2.x, s2 = r2.y, a2 = r2.z, bl = r2.w, pj = r3.x, kappa = r3.x + r3.y;
        float U[9], Th[9], Ah[9];
        rot_of(q, U);
        rotate_in(U, T, Th);
        const float tr = Th[0] + Th[4] + Th[8];
        Ah[0] = (a2 * s0 * s0 - pj) * Th[0] + kappa * tr;
        Ah[4] = (a2 * s1 * s1 - pj) * Th[4] + kappa * tr;
        Ah[8] = (a2 * s2 * s2 - pj) * Th[8] + kappa * tr;
        apply_pair(a2, bl, s0, s1, s2, Th[1], Th[3], Ah[1], Ah[3]);
        apply_pair(a2, bl, s0, s2, s1, Th[2], Th[6], Ah[2], Ah[6]);
        apply_pair(a2, bl, s1, s2, s0, Th[5], Th[7], Ah[5], Ah[7]);
        rotate_out(U, Ah, A);
    } else {
        const float* tri = rec + D;
#pragma unroll
        for (int r = 0; r < DD; ++r) A[r] = 0.0f;
#pragma unroll
        for (int r = 0; r < DD; ++r)
#pragma unroll
            for (int c = r; c < DD; ++c) {
                const float v = tri[op_index<D>(r, c)];
                A[r] += v * T[c];
                if (c != r) A[c] += v * T[r];
            }
    }
}

template <int D>
__device__ inline void store_operator_eos(const KP& p, int i, float c, const float* xo) {
    float rec[OpRec<D>::W];
#pragma unroll
    for (int k = 0; k < OpRec<D>::W; ++k) rec[k] = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a) rec[a] = xo[a];
    if constexpr (D == 3) {
        // U = I, S = (1, 1, 1), A = B = P = 0, Q = c: every entry of the
        // diagonal block c, empty pair blocks
        rec[3] = 1.0f;
        rec[7] = rec[8] = rec[9] = 1.0f;
        rec[13] = c;
    } else {
        rec[D + op_index<D>(0, 0)] = c;
        rec[D + op_index<D>(0, 3)] = c;
        rec[D + op_index<D>(3, 3)] = c;
    }
    float4* dst = reinterpret_cast<float4*>(p.pop + (int64_t)i * OpRec<D>::W);
#pragma unroll
    for (int k = 0; k < OpRec<D>::W / 4; ++k) dst[k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
}

// Separable stencil contractions for the CG operator.  With
// grad w = (dwx wy wz, wx dwy wz, wx wy dwz) the 3^D-node sums factor axis
// by axis: 270 multiply-adds per particle in 3D instead of ~400 for the
// node-by-node form.  `base` is the tile index of the particle's stencil
// corner (its cell - 1 in every axis).
template <int D>
__device__ inline void grad_gather(const float (*w)[3], const float (*dw)[3], int base,
                                   const float (*ut)[Dims<D>::LTN], float* T) {
    constexpr int LX = Dims<D>::LX, LY = Dims<D>::LY;
    if constexpr (D == 3) {
        float y00[3][3], y10[3][3], y01[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) y00[i][cc] = y10[i][cc] = y01[i][cc] = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int t0 = base + i * LX + j * LY;
                // the row's nine tile values load before its contractions
                float u[3][3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                    for (int k = 0; k < 3; ++k) u[cc][k] = ut[cc][t0 + k];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const float u0 = u[cc][0], u1 = u[cc][1], u2 = u[cc][2];
                    const float z0 = w[2][0] * u0 + w[2][1] * u1 + w[2][2] * u2;
                    const float z1 = dw[2][0] * u0 + dw[2][1] * u1 + dw[2][2] * u2;
                    y00[i][cc] += w[1][j] * z0;
                    y10[i][cc] += dw[1][j] * z0;
                    y01[i][cc] += w[1][j] * z1;
                }
            }
        }
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            T[0 * 3 + cc] = dw[0][0] * y00[0][cc] + dw[0][1] * y00[1][cc] + dw[0][2] * y00[2][cc];
            T[1 * 3 + cc] = w[0][0] * y10[0][cc] + w[0][1] * y10[1][cc] + w[0][2] * y10[2][cc];
            T[2 * 3 + cc] = w[0][0] * y01[0][cc] + w[0][1] * y01[1][cc] + w[0][2] * y01[2][cc];
        }
    } else {
        float z0[3][2], z1[3][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int t0 = base + i * LX;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const float u0 = ut[cc][t0], u1 = ut[cc][t0 + 1], u2 = ut[cc][t0 + 2];
                z0[i][cc] = w[1][0] * u0 + w[1][1] * u1 + w[1][2] * u2;
                z1[i][cc] = dw[1][0] * u0 + dw[1][1] * u1 + dw[1][2] * u2;
            }
        }
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            T[0 * 2 + cc] = dw[0][0] * z0[0][cc] + dw[0][1] * z0[1][cc] + dw[0][2] * z0[2][cc];
            T[1 * 2 + cc] = w[0][0] * z1[0][cc] + w[0][1] * z1[1][cc] + w[0][2] * z1[2][cc];
        }
    }
}

// q_node[cc] += sum_r A[r][cc] grad_r w_node over the stencil, into a
// per-wave LDS tile (channel-major, TN floats per channel).  Lanes of a wave
// hold distinct cells, so one stencil node is read-modified-written by at
// most one lane per step; the compiler fence keeps the steps in order.
template <int D>
__device__ inline void grad_scatter(const float (*w)[3], const float (*dw)[3], int base, const float* A, float* q) {
    constexpr int LX = Dims<D>::LX, LY = Dims<D>::LY, TN = Dims<D>::LTN;
    if constexpr (D == 3) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float x0[3], x1[3], x2[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                x0[cc] = A[0 * 3 + cc] * dw[0][i];
                x1[cc] = A[1 * 3 + cc] * w[0][i];
                x2[cc] = A[2 * 3 + cc] * w[0][i];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float ya[3], yb[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    ya[cc] = fmaf(w[1][j], x0[cc], dw[1][j] * x1[cc]);
                    yb[cc] = w[1][j] * x2[cc];
                }
                const int t0 = base + i * LX + j * LY;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    // a step's three reads go out before its writes
                    float o[3];
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) o[cc] = q[cc * TN + t0 + k];
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) q[cc * TN + t0 + k] = fmaf(w[2][k], ya[cc], fmaf(dw[2][k], yb[cc], o[cc]));
                    MPM_STEP_FENCE();
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float x0[2], x1[2];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                x0[cc] = A[0 * 2 + cc] * dw[0][i];
                x1[cc] = A[1 * 2 + cc] * w[0][i];
            }
            const int t0 = base + i * LX;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    float* dst = q + cc * TN + t0 + j;
                    *dst = fmaf(w[1][j], x0[cc], fmaf(dw[1][j], x1[cc], *dst));
                }
                MPM_STEP_FENCE();
            }
        }
    }
}

// ============================================================================
// P2G (Rasterize, MPM_Example.cpp:520-537): m_i += w m_p, (mv)_i += w m_p V_p.
// One wave per block (the waves of a work-group take different blocks and
// share nothing).  Lane c runs every particle of cell c, summing the cell's
// 3^D-node contribution in registers (the particles of a cell share its
// stencil), then adds it to the wave's LDS tile in 3^D ordered
// read-add-write steps (lanes hold distinct cells) and the wave stores the
// tile.  LDS traffic is one stencil per cell instead of one per particle.
// ============================================================================
template <int D>
__global__ __launch_bounds__(BWG) void k_p2g(KP p) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, C = D + 1, TN = Dm::TN, LX = Dm::LX, LY = Dm::LY;
    constexpr int NS = D == 3 ? 27 : 9;
    __shared__ float tile_all[NW][C * LTN];
    const int c = threadIdx.x & 63;
    float* mine = tile_all[threadIdx.x >> 6];
    int l[D];
    local_coords<D>(c, l);
    int base;  // tile index of this cell's stencil corner
    {
        int off0[D];
        base = tile_index<D>(l, 0, off0);
    }
    const XcdRange xr = xcd_wave_blocks(p.ctr->nblocks);
    for (int kb = xr.first; kb < xr.end; kb += xr.step) {
        // one scene: heavy blocks first within the XCD's range; batches keep
        // slot order, measured faster there (small blocks of many scenes, and
        // neighbouring waves share a scene in slot order)
        const int b = p.nscenes > 1 ? kb : p.block_order[kb];
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.active || block_empty(p, b)) continue;
        for (int t = c; t < C * LTN; t += 64) mine[t] = 0.0f;
        int bb[D];
        block_coords<D>(p, b, bb);
        float acc[NS][C];
#pragma unroll
        for (int n = 0; n < NS; ++n)
#pragma unroll
            for (int ch = 0; ch < C; ++ch) acc[n][ch] = 0.0f;
        const int i0 = p.cell_start[b * 65 + c], i1 = p.cell_start[b * 65 + c + 1];
        for (int i = i0; i < i1; ++i) {
            float w[D][3], dw[D][3], X[D];
            if (!cell_particle<D>(p, sc, i, bb, l, w, dw, X)) continue;
            float mv[C];
            mv[0] = p.pcur[PF<D>::M * p.pcap + i];
#pragma unroll
            for (int a = 0; a < D; ++a) mv[1 + a] = mv[0] * p.pcur[(PF<D>::V + a) * p.pcap + i];
            if constexpr (D == 3) {
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) {
                        const float wij = w[0][ii] * w[1][jj];
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk) {
                            const float wt = wij * w[2][kk];
#pragma unroll
                            for (int ch = 0; ch < C; ++ch) acc[(ii * 3 + jj) * 3 + kk][ch] += wt * mv[ch];
                        }
                    }
            } else {
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) {
                        const float wt = w[0][ii] * w[1][jj];
#pragma unroll
                        for (int ch = 0; ch < C; ++ch) acc[ii * 3 + jj][ch] += wt * mv[ch];
                    }
            }
        }
        MPM_STEP_FENCE();
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            const int t = D == 3 ? base + (n / 9) * LX + ((n / 3) % 3) * LY + n % 3 : base + (n / 3) * LX + n % 3;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) mine[ch * LTN + t] += acc[n][ch];
            MPM_STEP_FENCE();
        }
        float* dst = p.tiles + (int64_t)b * C * TN;
        for (int t = c; t < C * TN; t += 64) {
            const int cc = t / TN;
            int nb, local, tc[D];
            tile_map<D>(t - cc * TN, nb, local, tc);
            dst[t] = mine[cc * LTN + lds_index<D>(tc)];
        }
        MPM_STEP_FENCE();
    }
}

// Tile sum of P2G + Flag_Setup_Helper (Flag_Setup_Helper.h:25-35) +
// Velocity_Normalization_Helper (Velocity_Normalization_Helper.h:39-52)
template <int D>
__global__ __launch_bounds__(WG) void k_normalize(KP p) {
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    for (int b = r.first; b < r.end; b += r.step) {
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.active) continue;
        const int n = b * 64 + lane;
        float v[D + 1];
        gather_node_tiles<D, D + 1>(p, sc, n, v);
        const float m = v[0];
        const bool interior = m > 0.0f;
        p.g_m[n] = m;
        p.g_flags[n] = interior ? MPM_FLAG_INTERIOR_DEV : 0u;
        const float inv = interior ? 1.0f / m : 0.0f;
#pragma unroll
        for (int a = 0; a < D; ++a) p.g_v[a * p.gcap + n] = interior ? v[1 + a] * inv : v[1 + a];
    }
}

// gather of `nch` grid channel arrays into an LDS tile
template <int D, int NCH>
__device__ inline void gather_tile(const KP& p, const int* nbr, float (*tile)[Dims<D>::LTN], const float* const* src) {
    constexpr int TN = Dims<D>::TN;
    for (int t = threadIdx.x; t < TN; t += (int)blockDim.x) {
        int nb, local, tc[D];
        tile_map<D>(t, nb, local, tc);
        const int blk = nbr[nb];
        const int li = lds_index<D>(tc);
#pragma unroll
        for (int c = 0; c < NCH; ++c) tile[c][li] = blk >= 0 ? src[c][(int64_t)blk * 64 + local] : 0.0f;
    }
}

// Process_Waiting_Particles (MPM_Example.cpp:827-840): V0 = m / (sum w m_i / dx^d)
template <int D>
__global__ __launch_bounds__(BWG) void k_init_volume(KP p) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, NBR = Dm::NBR;
    __shared__ float tile[1][LTN];
    __shared__ int nbr[NBR];
    const int nblocks = p.ctr->nblocks;
    const XcdRange xr = xcd_blocks(nblocks);
    for (int kb = xr.first; kb < xr.end; kb += xr.step) {
        const int b = p.block_order[kb];  // heavy blocks first within the XCD's range
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.waiting || block_empty(p, b)) continue;
        float vol = 1.0f;
        for (int a = 0; a < D; ++a) vol *= sc.dx;
        const float inv_vol = 1.0f / vol;
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        block_coords<D>(p, b, bb);
        __syncthreads();
        const float* src[1] = {p.g_m};
        gather_tile<D, 1>(p, nbr, tile, src);
        __syncthreads();
        const int i1 = p.block_pstart[b + 1];
        for (int i = p.block_pstart[b] + (int)threadIdx.x; i < i1; i += BWG) {
            // the waiting particles are those without a volume yet (new or emitted)
            if (p.pcur[PF<D>::VOL * p.pcap + i] != 0.0f) continue;
            float w[D][3], dw[D][3], X[D];
            int l[D];
            if (!particle_weights<D>(p, sc, i, bb, w, dw, l, X)) continue;
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
// Explicit force of one particle into its wave's tile:
// f_i += g m w_i - A grad w_i, A = V0 P F^T (MPM_Example.cpp:774-784).
template <int D>
__device__ inline void scatter_force(const SceneParams& sc, const int* l, const float (*w)[3], const float (*dw)[3],
                                     const float* A, float m, float* mine) {
    constexpr int LTN = Dims<D>::LTN, ST = Dims<D>::ST;
    float gm[D];
#pragma unroll
    for (int a = 0; a < D; ++a) gm[a] = sc.gravity[a] * m;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
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
            mine[a * LTN + t] += gm[a] * wt - inner;
        }
        MPM_STEP_FENCE();
    }
}

// Constitutive update (MPM_Example.cpp:563-571) of particle i of material
// mat (CONST: plasticity, F, Jp and the CG operator record) and its stress
// term A of the explicit force (FORCE: f_i -= A grad w_i with A = V0 P F^T,
// MPM_Example.cpp:757-784); an EOS particle has no constitutive model and its
// term is A = -(V0/rho) K rho^(gamma-1) I (MPM_Example.cpp:774-782).
// ============================================================================
// The operator record's position: particle i's stencil offsets, the same
// arithmetic as the weights of its binned cell (binning and the record are
// formed from the same position, and particles do not move between them).
template <int D>
__device__ inline void particle_offsets(const KP& p, const SceneParams& sc, int i, float* xo) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
        int c;
        xo[a] = stencil_offset(p.pcur[(PF<D>::X + a) * p.pcap + i], sc.dx, sc.inv_dx, c);
    }
}

template <int D, bool CONST, bool FORCE>
__device__ __forceinline__ void particle_stress(const KP& p, const SceneParams& sc, const Material& mat, int i,
                                                float* A) {
    constexpr int DD = D * D;
    float Jp = p.pcur[PF<D>::J * p.pcap + i];
    const float V0 = p.pcur[PF<D>::VOL * p.pcap + i];
    if (mat.eos) {
        // EOS fluid (Jp slot = density): no constitutive model
        const float rg = powf(Jp, mat.gamma - 1.0f);
        if (CONST) {
            float xo[D];
            particle_offsets<D>(p, sc, i, xo);
            store_operator_eos<D>(p, i, V0 * mat.K * mat.gamma * rg, xo);
            if (!FORCE) {
#pragma unroll
                for (int k = 0; k < DD; ++k) p.rcache[k * p.pcap + i] = (k % (D + 1) == 0) ? 1.0f : 0.0f;
            }
        }
        // f += (V0/rho) K rho^(gamma-1) grad w
        const float coeff = V0 / Jp * mat.K * rg;
#pragma unroll
        for (int k = 0; k < DD; ++k) A[k] = (k % (D + 1) == 0) ? -coeff : 0.0f;
    } else {
        float F[DD], R[DD];
#pragma unroll
        for (int k = 0; k < DD; ++k) F[k] = p.pcur[(PF<D>::F + k) * p.pcap + i];
        if (CONST) {
            float U[DD];
            double Sv[D];
            plasticity<D>(F, Jp, R, mat, U, Sv);
#pragma unroll
            for (int k = 0; k < DD; ++k) {
                p.pcur[(PF<D>::F + k) * p.pcap + i] = F[k];
                // R is read back only by the staged explicit force
                if (!FORCE) p.rcache[k * p.pcap + i] = R[k];
            }
            p.pcur[PF<D>::J * p.pcap + i] = Jp;
            StressState<D> so;
            stress_state<D>(F, R, Jp, mat, so);
            float xo[D];
            particle_offsets<D>(p, sc, i, xo);
            store_operator<D>(p, i, so, V0, xo, U, Sv);
        } else {
#pragma unroll
            for (int k = 0; k < DD; ++k) R[k] = p.rcache[k * p.pcap + i];
        }
        if (FORCE) {
            StressState<D> st;
            stress_state<D>(F, R, Jp, mat, st);
            float P[DD];
            first_pk<D>(st, P);
            // V0 P F^T
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int cc = 0; cc < D; ++cc) {
                    float s = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) s += P[r * D + k] * F[cc * D + k];
                    A[r * D + cc] = s * V0;
                }
        }
    }
}

// One work-group per block, its particles in chunks of PF_CHUNK (a full
// block of 8 per cell is one chunk).  Each chunk is first updated a particle
// per thread in storage order (coalesced loads and stores; a wave past the
// block's last particle skips its part), one pass per entry of the scene's
// material table, each pass taking the particles of that entry (a single
// pass without a check in a one-material scene): the material's parameters
// stay scalar operands.  With FORCE the chunk's stress terms go to LDS and
// are then scattered cell by cell (lane = cell, waves interleave a cell's
// particles) into the wave tiles.
constexpr int PF_CHUNK = 2 * BWG;
template <int D, bool CONST, bool FORCE>
__global__ __launch_bounds__(BWG) void k_particle_force(KP p) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, DD = D * D;
    __shared__ __align__(16) float tile[NW * D * LTN];
    __shared__ float stress[DD][PF_CHUNK];
    const int nblocks = p.ctr->nblocks;
    const int c = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* mine = tile + wv * D * LTN;
    int l[D];
    local_coords<D>(c, l);
    const XcdRange xr = xcd_blocks(nblocks);
    for (int kb = xr.first; kb < xr.end; kb += xr.step) {
        const int b = p.block_order[kb];  // heavy blocks first within the XCD's range
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.active || block_empty(p, b)) continue;
        int bb[D];
        slot_coords<D>(p.block_slot[b] - sc.slot_base, sc.nb, bb);  // computed: the loaded form costs occupancy here
        if (FORCE)
            for (int t = threadIdx.x; t < NW * D * LTN; t += BWG) tile[t] = 0.0f;
        const int p0 = p.block_pstart[b], p1 = p.block_pstart[b + 1];
        const int c0 = p.cell_start[b * 65 + c], c1 = p.cell_start[b * 65 + c + 1];
        const int nmat = sc.nmat;
        for (int chunk = p0; chunk < p1; chunk += PF_CHUNK) {
            for (int half = 0; half < PF_CHUNK; half += BWG) {
                const int t = half + (int)threadIdx.x, i = chunk + t;
                if (chunk + half + (int)(threadIdx.x & ~63u) >= p1) continue;  // the wave's part is past the block
                for (int u = 0; u < nmat; ++u) {
                    if (i < p1 && (nmat == 1 || __float_as_int(p.pcur[PF<D>::MAT * p.pcap + i]) == u)) {
                        float A[DD];
                        particle_stress<D, CONST, FORCE>(p, sc, sc.mats[u], i, A);
                        if (FORCE) {
#pragma unroll
                            for (int k = 0; k < DD; ++k) stress[k][t] = A[k];
                        }
                    }
                }
            }
            if (FORCE) {
                __syncthreads();  // the chunk's stress terms (and, first, the zeroed tiles)
                const int lo = max(c0, chunk), hi = min(c1, chunk + PF_CHUNK);
                for (int j = lo + wv; j < hi; j += NW) {
                    float w[D][3], dw[D][3], X[D];
                    if (!cell_particle<D>(p, sc, j, bb, l, w, dw, X)) continue;
                    float A[DD];
#pragma unroll
                    for (int k = 0; k < DD; ++k) A[k] = stress[k][j - chunk];
                    scatter_force<D>(sc, l, w, dw, A, p.pcur[PF<D>::M * p.pcap + j], mine);
                }
                __syncthreads();  // before the next chunk overwrites the stress terms
            }
        }
        if (FORCE) {
            store_wave_tiles<D, D>(p, b, tile);
            __syncthreads();
        }
    }
}

// ============================================================================
// Grid update: tile sum of the force, Explicit_Force_Helper (v* = v + dt f/m),
// plane collision (Grid_Based_Collision_Helper.h:57-85) and the CG set-up
// (MPM_RHS_Helper, Reset_Solver_Channels: r = b = v*, s = 0).
// WRITE_F keeps the force in g_q for inspection.
// ============================================================================
template <int D, bool FORCE, bool COLL, bool CGINIT, bool WRITE_F>
__global__ __launch_bounds__(WG) void k_grid_update(KP p) {
    if (CGINIT) {
        // a new solve for every scene (inactive scenes finish at the first
        // test; on the single-scene path the others are finished here)
        for (int si = blockIdx.x * WG + threadIdx.x; si < p.nscenes; si += gridDim.x * WG) {
            CGState& cg = p.cg[si];
            const int other = p.cg_target == 1 && si != p.solo ? 1 : 0;
            cg.iter = 0;
            cg.done = other;
            cg.converged = other;
            cg.multiplies = 0;
            cg.rho_hist[0] = cg.rho_hist[1] = 0.0;
            cg.alpha = cg.beta = cg.conv = cg.rr = cg.sq = 0.0;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            p.ctr->cg_ndone = 0;
            p.ctr->cg_iters = 0;
        }
    }
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    for (int b = r.first; b < r.end; b += r.step) {
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.active) continue;
        const int n = b * 64 + lane;
        const float dt = sc.dt;
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float vs[D];
        if (FORCE) {
            float f[D];
            gather_node_tiles<D, D>(p, sc, n, f);
            if (WRITE_F) {
#pragma unroll
                for (int a = 0; a < D; ++a) p.g_q[a * p.gcap + n] = f[a];
            }
            if (interior) {
                const float dtm = dt / p.g_m[n];
#pragma unroll
                for (int a = 0; a < D; ++a) vs[a] = p.g_v[a * p.gcap + n] + dtm * f[a];
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
            node_center<D>(p, sc, n, x);
            for (int id = 0; id < sc.nplanes; ++id) {
                if (!inside_plane(sc, id, x, D)) continue;
                float proj = 0.0f;
#pragma unroll
                for (int a = 0; a < D; ++a) proj += vs[a] * sc.pn[id][a];
                if (proj < 0.0f) {
#pragma unroll
                    for (int a = 0; a < D; ++a) vs[a] -= sc.pn[id][a] * proj;
                    float nrm = 0.0f;
#pragma unroll
                    for (int a = 0; a < D; ++a) nrm += vs[a] * vs[a];
                    nrm = sqrtf(nrm);
                    const float mu = sc.pmu[id];
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
            }
        }
    }
}

// ============================================================================
// CG (Apply_Force's solve, MPM_Example.cpp:700-711) for every scene of the
// batch at once: Hestenes-Stiefel CG in the M inner product with a restart
// at iteration 0.  Launch order per solve:
//   grid_update<CGINIT>, cg_force<RESTART>, cg_restart_node, cg_reduce_rr(0),
//   then for k = 0, 1, ...: cg_force(k), cg_node1(k), cg_reduce_sq(k),
//   cg_node2, cg_reduce_rr(k + 1).
// A scene whose solve is finished is skipped by every kernel; when all are,
// the kernels return at once.
// ============================================================================

// s = r (k = 0) or r + beta s: the one formula shared by the operator's
// on-the-fly gather and node1's store, so both see identical bits
__device__ inline float cg_direction(float r, float s, float beta, int k) {
    return k == 0 ? r : fmaf(beta, s, r);
}

__device__ inline bool cg_all_done(const KP& p) { return p.ctr->cg_ndone >= p.cg_target; }

// Single-scene CG kernels in a batch or arena (one active scene): blocks of
// the other scenes are skipped.
__device__ inline bool other_scene(const KP& p, int b) { return p.nscenes > 1 && p.block_scene[b] != p.solo; }

// MPM_CG_System::Force (MPM_CG_System.h:150-215) with the linearised operator
// record: T = sum_i grad w_ip (x) u_i;  A = L T;  q_i += A^T grad w_ip.
// RESTART applies it to v* (the initial guess); otherwise to the search
// direction of iteration k, s = r + beta s, formed while gathering.
template <int D, bool RESTART, bool ONE>
__global__ __launch_bounds__(BWG) void k_cg_force(KP p, int k) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, NBR = Dm::NBR, DD = D * D;
    __shared__ float ut[D][LTN];
    __shared__ __align__(16) float qt[NW * D * LTN];
    __shared__ int nbr[NBR];
    if (cg_all_done(p)) return;
    const int nblocks = p.ctr->nblocks;
    const int c = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* mine = qt + wv * D * LTN;
    TileMaps<D> tm;
    tm.init();
    int l[D];
    local_coords<D>(c, l);
    int base;  // tile index of this cell's stencil corner
    {
        int off0[D];
        base = tile_index<D>(l, 0, off0);
    }
    const XcdRange xr = xcd_blocks(nblocks);
    for (int kb = xr.first; kb < xr.end; kb += xr.step) {
        const int b = p.block_order[p.bcap + kb];  // the fullest cells first within the XCD's range
        // one scene: no per-block lookups (its solve is live: cg_all_done above)
        if (ONE && other_scene(p, b)) continue;
        const int si = ONE ? p.solo : p.block_scene[b];
        const SceneParams& sc = p.scn[si];
        if (!sc.active || (!ONE && p.cg[si].done) || block_empty(p, b)) continue;
        const float beta = (RESTART || ONE) ? 0.0f : (float)p.cg[si].beta;
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        static_assert((NW * D * LTN) % 4 == 0, "float4 zeroing");
        for (int t = threadIdx.x; t < NW * D * LTN / 4; t += BWG)
            reinterpret_cast<float4*>(qt)[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
        if (tm.g != ~0u) {
            const int blk = nbr[tm.g & 31u];
            const int li = (int)(tm.g >> 11);
            const int64_t n = (int64_t)blk * 64 + (int)((tm.g >> 5) & 63u);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                float u = 0.0f;
                if (blk >= 0)
                    u = RESTART ? p.g_vs[a * p.gcap + n]
                        : ONE   ? p.g_s[a * p.gcap + n]
                                : cg_direction(p.g_r[a * p.gcap + n], p.g_s[a * p.gcap + n], beta, k);
                ut[a][li] = u;
            }
        }
        __syncthreads();
        const int i0 = p.cell_start[b * 65 + c], i1 = p.cell_start[b * 65 + c + 1];
        for (int i = i0 + wv; i < i1; i += NW) {
            const float* rec = p.pop + (int64_t)i * OpRec<D>::W;
            float w[D][3], dw[D][3];
            {
                // the record's stencil offsets, formed from the position that
                // binned the particle into this cell
                const float4 x4 = *reinterpret_cast<const float4*>(rec);
                const float xo[3] = {x4.x, x4.y, x4.z};
#pragma unroll
                for (int a = 0; a < D; ++a) bspline_from_offset(xo[a], sc.inv_dx, w[a], dw[a]);
            }
            float T[DD];
            grad_gather<D>(w, dw, base, ut, T);
            float A[DD];  // V0 F dP(T^T F)^T
            apply_operator<D>(rec, T, A);
            grad_scatter<D>(w, dw, base, A, mine);
        }
        __syncthreads();
        store_wave_tiles_mapped<D>(p, b, qt, tm.s);
        __syncthreads();
    }
}

// The operator's scatter for one particle, accumulated in registers over the
// particles of a cell (they share the cell's stencil): acc[node * D + cc],
// node = (i * 3 + j) * 3 + k.
template <int D>
__device__ inline void grad_accumulate(const float (*w)[3], const float (*dw)[3], const float* A, float* acc) {
    if constexpr (D == 3) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float x0[3], x1[3], x2[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                x0[cc] = A[0 * 3 + cc] * dw[0][i];
                x1[cc] = A[1 * 3 + cc] * w[0][i];
                x2[cc] = A[2 * 3 + cc] * w[0][i];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float ya[3], yb[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    ya[cc] = fmaf(w[1][j], x0[cc], dw[1][j] * x1[cc]);
                    yb[cc] = w[1][j] * x2[cc];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        float& a = acc[((i * 3 + j) * 3 + k) * 3 + cc];
                        a = fmaf(w[2][k], ya[cc], fmaf(dw[2][k], yb[cc], a));
                    }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float x0[2], x1[2];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                x0[cc] = A[0 * 2 + cc] * dw[0][i];
                x1[cc] = A[1 * 2 + cc] * w[0][i];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    float& a = acc[(i * 3 + j) * 2 + cc];
                    a = fmaf(w[1][j], x0[cc], fmaf(dw[1][j], x1[cc], a));
                }
        }
    }
}

// Adds a lane's accumulated stencil to the wave's LDS tile: one ordered
// read-add-write per stencil node (lanes hold distinct cells).
template <int D>
__device__ inline void flush_accumulated(int base, const float* acc, float* q) {
    constexpr int LX = Dims<D>::LX, LY = Dims<D>::LY, TN = Dims<D>::LTN;
    if constexpr (D == 3) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int t = base + i * LX + j * LY + k;
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) q[cc * TN + t] += acc[((i * 3 + j) * 3 + k) * 3 + cc];
                    MPM_STEP_FENCE();
                }
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int t = base + i * LX + j;
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) q[cc * TN + t] += acc[(i * 3 + j) * 2 + cc];
                MPM_STEP_FENCE();
            }
    }
}

// The same operator with one wave per block: lane c runs every particle of
// cell c and accumulates the cell's scatter in registers, so a block costs
// one read-add-write chain per cell instead of one per particle.  The waves
// of a work-group take different blocks and share nothing (no barriers).
template <int D, bool RESTART, bool ONE>
__global__ __launch_bounds__(WG) void k_cg_force_cell(KP p, int k) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, NBR = Dm::NBR, DD = D * D, TN = Dm::TN, W = WG / 64;
    __shared__ float ut_all[W][D * LTN];
    __shared__ float qt_all[W][D * LTN];
    __shared__ int nbr_all[W][NBR];
    if (cg_all_done(p)) return;
    const int c = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* ut = ut_all[wv];
    float* qt = qt_all[wv];
    int* nbr = nbr_all[wv];
    int l[D];
    local_coords<D>(c, l);
    int base;
    {
        int off0[D];
        base = tile_index<D>(l, 0, off0);
    }
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    for (int b = r.first; b < r.end; b += r.step) {
        if (ONE && other_scene(p, b)) continue;
        const int si = ONE ? p.solo : p.block_scene[b];
        const SceneParams& sc = p.scn[si];
        if (!sc.active || (!ONE && p.cg[si].done) || block_empty(p, b)) continue;
        const float beta = (RESTART || ONE) ? 0.0f : (float)p.cg[si].beta;
        if (c < NBR) nbr[c] = p.block_nbr[b * NBR + c];
        for (int t = c; t < D * LTN; t += 64) qt[t] = 0.0f;
        MPM_STEP_FENCE();
        for (int t = c; t < TN; t += 64) {
            int nb, local, tc[D];
            tile_map<D>(t, nb, local, tc);
            const int blk = nbr[nb];
            const int li = lds_index<D>(tc);
            const int64_t n = (int64_t)blk * 64 + local;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                float u = 0.0f;
                if (blk >= 0)
                    u = RESTART ? p.g_vs[a * p.gcap + n]
                        : ONE   ? p.g_s[a * p.gcap + n]
                                : cg_direction(p.g_r[a * p.gcap + n], p.g_s[a * p.gcap + n], beta, k);
                ut[a * LTN + li] = u;
            }
        }
        MPM_STEP_FENCE();
        float acc[D == 3 ? 81 : 18];
#pragma unroll
        for (int t = 0; t < (D == 3 ? 81 : 18); ++t) acc[t] = 0.0f;
        const int i0 = p.cell_start[b * 65 + c], i1 = p.cell_start[b * 65 + c + 1];
        const float (*utile)[LTN] = reinterpret_cast<const float (*)[LTN]>(ut);
        for (int i = i0; i < i1; ++i) {
            const float* rec = p.pop + (int64_t)i * OpRec<D>::W;
            float w[D][3], dw[D][3];
            {
                // the record's stencil offsets, formed from the position that
                // binned the particle into this cell
                const float4 x4 = *reinterpret_cast<const float4*>(rec);
                const float xo[3] = {x4.x, x4.y, x4.z};
#pragma unroll
                for (int a = 0; a < D; ++a) bspline_from_offset(xo[a], sc.inv_dx, w[a], dw[a]);
            }
            float T[DD];
            grad_gather<D>(w, dw, base, utile, T);
            float A[DD];
            apply_operator<D>(rec, T, A);
            grad_accumulate<D>(w, dw, A, acc);
        }
        flush_accumulated<D>(base, acc, qt);
        MPM_STEP_FENCE();
        float* dst = p.tiles + (int64_t)b * D * TN;
        for (int t = c; t < D * TN; t += 64) {
            const int cc = t / TN;
            int nb, local, tc[D];
            tile_map<D>(t - cc * TN, nb, local, tc);
            dst[t] = qt[cc * LTN + lds_index<D>(tc)];
        }
        MPM_STEP_FENCE();
    }
}

// Project (Project_Helper.h:30-45): zero at nodes inside a plane
template <int D>
__device__ inline bool project_zero(const KP& p, const SceneParams& sc, int n) {
    if (sc.project_mode != 0 || sc.nplanes == 0) return false;
    float x[D];
    node_center<D>(p, sc, n, x);
    for (int id = 0; id < sc.nplanes; ++id)
        if (inside_plane(sc, id, x, D)) return true;
    return false;
}

// sum and max over a wave; result valid in every lane
__device__ inline void wave_reduce2(double& sum, double& mx) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        mx = fmax(mx, __shfl_xor(mx, off));
    }
}

// Two reduction layouts.  Batches (ONE = false): node kernels write one
// partial per block and a one-work-group-per-scene kernel reduces each
// scene's partials.  A batch of one scene (ONE = true): node kernels write
// one partial per work-group and every work-group of the consuming kernel
// reduces all of them itself (same values, same order, same result), which
// saves the two reducer launches per CG iteration.

// Writes this wave's partials (ONE: accumulated over its blocks) -- per block
// in batches, per work-group for one scene.
__device__ inline void wg_partials_store(double* dsum, double* dmax, double sum, double mx) {
    __shared__ double ws[WG / 64], wm[WG / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { ws[wave] = sum; wm[wave] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, m = 0.0;
#pragma unroll
        for (int w = 0; w < WG / 64; ++w) { s += ws[w]; m = fmax(m, wm[w]); }
        dsum[blockIdx.x] = s;
        if (dmax) dmax[blockIdx.x] = m;
    }
}

// Sum / max of `count` partials starting at `first`, in a fixed order
// relative to `first`; the result is broadcast to every thread.
__device__ inline void reduce_range(const double* sums, const double* maxs, int first, int count, double& sum,
                                    double& mx) {
    constexpr int U = 8;  // eight independent loads in flight per thread
    double sa[U], ma[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { sa[u] = 0.0; ma[u] = 0.0; }
    for (int base = (int)threadIdx.x; base < count; base += U * WG) {
        // the chunk's loads all go out before its first add (an index past
        // the range reads the range's last entry and is not added)
        double vs[U], vm[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = min(base + u * WG, count - 1);
            vs[u] = sums[first + j];
            vm[u] = maxs ? maxs[first + j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * WG < count) {
                sa[u] += vs[u];
                if (maxs) ma[u] = fmax(ma[u], vm[u]);
            }
    }
    double s = 0.0, m = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) { s += sa[u]; m = fmax(m, ma[u]); }
    wave_reduce2(s, m);
    __shared__ double ss[WG / 64], sm[WG / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) { ss[wave] = s; sm[wave] = m; }
    __syncthreads();
    s = 0.0;
    m = 0.0;
#pragma unroll
    for (int w = 0; w < WG / 64; ++w) { s += ss[w]; m = fmax(m, sm[w]); }
    sum = s;
    mx = m;
}

// Convergence test of iteration k (Convergence_Norm, MPM_CG_System.h:241-249:
// L-inf over nodes of |r|) on r.r and max |r|^2, and beta = rho_k / rho_{k-1}.
// A scene is done when the test passes (after cg_min iterations) or at its
// cap cg_max.  `record`: this thread writes the scene's state.
// The scene's inputs of the test, read in one go (before the partial
// reduction, so that their loads overlap it).
struct CGTestIn {
    double rho_prev, tol;
    int cg_min, cg_max;
    bool empty;
};
__device__ inline CGTestIn cg_test_in(const KP& p, int si, int k) {
    const SceneParams& sc = p.scn[si];
    CGTestIn in;
    in.rho_prev = p.cg[si].rho_hist[(k + 1) & 1];
    in.tol = sc.cg_tol;
    in.cg_min = sc.cg_min;
    in.cg_max = sc.cg_max;
    const int active = sc.active, b0 = p.sst[si].block_begin, b1 = p.sst[si].block_end;
    in.empty = (active == 0) | (b1 <= b0);
    return in;
}

__device__ inline bool cg_test(const KP& p, int si, int k, double rr, double mx, bool record, float& beta_f,
                               const CGTestIn& in) {
    CGState& cg = p.cg[si];
    const double conv = sqrt(mx);
    const bool pass = conv <= in.tol && (k >= in.cg_min || conv < (double)FLT_EPSILON);
    const bool done = pass || k >= in.cg_max || in.empty;
    const double beta = k == 0 ? 0.0 : rr / in.rho_prev;
    if (record) {
        if (k == 0) cg.multiplies = 1;  // the restart Multiply
        cg.iter = k;
        cg.conv = conv;
        cg.rr = rr;
        if (done) {
            cg.converged = (pass || in.empty) ? 1 : 0;
            cg.done = 1;
            atomicMax(&p.ctr->cg_iters, k);
            atomicAdd(&p.ctr->cg_ndone, 1);
        } else {
            cg.beta = beta;
            cg.rho_hist[k & 1] = rr;
        }
    }
    beta_f = (float)beta;
    return done;
}

// CG restart: q = dt^2/m q + x (Multiply_Helper), r = b - q, Project(r);
// partials of r.r (M-weighted) and max |r|^2.
template <int D, bool ONE>
__global__ __launch_bounds__(WG) void k_cg_restart_node(KP p) {
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    double wsum = 0.0, wmx = 0.0;
    for (int b = r.first; b < r.end; b += r.step) {
        if (ONE && other_scene(p, b)) continue;
        const SceneParams& sc = p.scn[ONE ? p.solo : p.block_scene[b]];
        if (!sc.active) continue;
        const int n = b * 64 + lane;
        const float sds = sc.dt * sc.dt;
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float q[D], rv[D];
        gather_node_tiles<D, D>(p, sc, n, q);
#pragma unroll
        for (int a = 0; a < D; ++a) rv[a] = p.g_r[a * p.gcap + n];
        if (interior) {
            const float cm = sds / p.g_m[n];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                q[a] = cm * q[a] + p.g_vs[a * p.gcap + n];
                rv[a] = rv[a] - q[a];
            }
        }
        if (project_zero<D>(p, sc, n)) {
#pragma unroll
            for (int a = 0; a < D; ++a) rv[a] = 0.0f;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            p.g_r[a * p.gcap + n] = rv[a];
            p.g_q[a * p.gcap + n] = q[a];
        }
        double sum = 0.0, mx = 0.0;
        if (interior) {
            const double m = (double)p.g_m[n];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                sum += (double)rv[a] * ((double)rv[a] * m);
                mx += (double)rv[a] * (double)rv[a];
            }
        }
        wave_reduce2(sum, mx);
        if (ONE) {
            wsum += sum;
            wmx = fmax(wmx, mx);
        } else if (lane == 0) {
            p.bsum[b] = sum;
            p.bmax[b] = mx;
        }
    }
    if (ONE) wg_partials_store(p.bsum, p.bmax, wsum, wmx);
}

// One scene: convergence test of iteration k on the r.r partials (every
// work-group reduces them; work-group 0 records the state), then
// s = r (k = 0) or r + beta s on the interior nodes.
template <int D>
__global__ __launch_bounds__(WG) void k_cg_update_s(KP p, int k, int nparts) {
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    // the wave's first block: r and s load while the partials are reduced
    // (a wave without blocks loads block 0's nodes and uses nothing)
    float r0[D], s0[D];
    const bool has = r.first < r.end;
    const bool mine = has && !other_scene(p, r.first);
    const int n0 = (has ? r.first : 0) * 64 + lane;
    const unsigned flags0 = p.g_flags[n0];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        r0[a] = p.g_r[a * p.gcap + n0];
        s0[a] = p.g_s[a * p.gcap + n0];
    }
    if (cg_all_done(p)) return;
    // the test's inputs load alongside the partials
    const CGTestIn tin = cg_test_in(p, p.solo, k);
    double rr, mx;
    reduce_range(p.bsum, p.bmax, 0, nparts, rr, mx);
    float beta;
    if (cg_test(p, p.solo, k, rr, mx, blockIdx.x == 0 && threadIdx.x == 0, beta, tin)) return;
    if (mine && (flags0 & MPM_FLAG_INTERIOR_DEV) != 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) p.g_s[a * p.gcap + n0] = cg_direction(r0[a], s0[a], beta, k);
    }
    for (int b = r.first + r.step; b < r.end; b += r.step) {
        if (other_scene(p, b)) continue;
        const int n = b * 64 + lane;
        if (!(p.g_flags[n] & MPM_FLAG_INTERIOR_DEV)) continue;
#pragma unroll
        for (int a = 0; a < D; ++a)
            p.g_s[a * p.gcap + n] = cg_direction(p.g_r[a * p.gcap + n], p.g_s[a * p.gcap + n], beta, k);
    }
}

// Tile sum of the operator, q = dt^2/m q + s, Project(q); partials of s.q.
// Batches form and store this iteration's s = r + beta s (interior nodes);
// for one scene k_cg_update_s stored it already.
template <int D, bool ONE>
__global__ __launch_bounds__(WG) void k_cg_node1(KP p, int k) {
    if (cg_all_done(p)) return;
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    double wsum = 0.0;
    for (int b = r.first; b < r.end; b += r.step) {
        if (ONE && other_scene(p, b)) continue;
        const int si = ONE ? p.solo : p.block_scene[b];
        const SceneParams& sc = p.scn[si];
        if (!sc.active || (!ONE && p.cg[si].done)) continue;
        const float beta = ONE ? 0.0f : (float)p.cg[si].beta;
        const float sds = sc.dt * sc.dt;
        const int n = b * 64 + lane;
        const bool interior = (p.g_flags[n] & MPM_FLAG_INTERIOR_DEV) != 0;
        float q[D], sv[D], mn = 1.0f;
        // one scene: s and m load alongside the flags and the tile gather
        // (every node of an active block has them; only interior ones use them)
        if (ONE) {
            mn = p.g_m[n];
#pragma unroll
            for (int a = 0; a < D; ++a) sv[a] = p.g_s[a * p.gcap + n];
        }
        gather_node_tiles<D, D>(p, sc, n, q);
        if (interior) {
            const float cm = sds / (ONE ? mn : p.g_m[n]);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (!ONE) {
                    sv[a] = cg_direction(p.g_r[a * p.gcap + n], p.g_s[a * p.gcap + n], beta, k);
                    p.g_s[a * p.gcap + n] = sv[a];
                }
                q[a] = cm * q[a] + sv[a];
            }
        }
        if (project_zero<D>(p, sc, n)) {
#pragma unroll
            for (int a = 0; a < D; ++a) q[a] = 0.0f;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) p.g_q[a * p.gcap + n] = q[a];
        double sum = 0.0, mx = 0.0;
        if (interior) {
            const double m = (double)(ONE ? mn : p.g_m[n]);
#pragma unroll
            for (int a = 0; a < D; ++a) sum += (double)sv[a] * ((double)q[a] * m);
        }
        wave_reduce2(sum, mx);
        if (ONE) wsum += sum;
        else if (lane == 0) p.bsq[b] = sum;
    }
    if (ONE) wg_partials_store(p.bsq, nullptr, wsum, 0.0);
}

// alpha = rho_k / s.q;  x += alpha s, r -= alpha q; partials of r.r and
// max |r|^2.  One scene: every work-group reduces the s.q partials itself
// (work-group 0 records alpha); batches read alpha from k_cg_reduce_sq.
template <int D, bool ONE>
__global__ __launch_bounds__(WG) void k_cg_node2(KP p, int k, int nparts) {
    const XcdRange r = xcd_wave_blocks(p.ctr->nblocks);
    const int lane = threadIdx.x & 63;
    // one scene: the wave's first block's node values load while the s.q
    // partials are reduced
    float pre_s[D], pre_q[D], pre_x[D], pre_r[D], pre_m = 0.0f;
    unsigned pre_flags = 0u;
    if (ONE && r.first < r.end) {
        const int n = r.first * 64 + lane;
        pre_flags = p.g_flags[n];
        pre_m = p.g_m[n];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            pre_s[a] = p.g_s[a * p.gcap + n];
            pre_q[a] = p.g_q[a * p.gcap + n];
            pre_x[a] = p.g_vs[a * p.gcap + n];
            pre_r[a] = p.g_r[a * p.gcap + n];
        }
    }
    if (cg_all_done(p)) return;
    float alpha1 = 0.0f;
    if (ONE) {
        // the scene's done flag and rho load alongside the partials (a
        // finished solve writes nothing)
        CGState& cg = p.cg[p.solo];
        const bool stop = cg.done != 0;
        const double rho = cg.rho_hist[k & 1];
        double sq, unused;
        reduce_range(p.bsq, nullptr, 0, nparts, sq, unused);
        if (stop) return;
        const double alpha_d = sq != 0.0 ? rho / sq : (double)FLT_MAX;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            cg.sq = sq;
            cg.alpha = alpha_d;
            cg.multiplies += 1;
        }
        alpha1 = (float)alpha_d;
    }
    double wsum = 0.0, wmx = 0.0;
    for (int b = r.first; b < r.end; b += r.step) {
        if (ONE && other_scene(p, b)) continue;
        const int si = ONE ? p.solo : p.block_scene[b];
        const SceneParams& sc = p.scn[si];
        if (!sc.active || (!ONE && p.cg[si].done)) continue;
        const float alpha = ONE ? alpha1 : (float)p.cg[si].alpha;
        const int n = b * 64 + lane;
        const bool pre = ONE && b == r.first;
        const bool interior = ((pre ? pre_flags : p.g_flags[n]) & MPM_FLAG_INTERIOR_DEV) != 0;
        double sum = 0.0, mx = 0.0;
        if (interior) {
            const double m = (double)(pre ? pre_m : p.g_m[n]);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float sv = pre ? pre_s[a] : p.g_s[a * p.gcap + n];
                const float q = pre ? pre_q[a] : p.g_q[a * p.gcap + n];
                const float xv = pre ? pre_x[a] : p.g_vs[a * p.gcap + n];
                p.g_vs[a * p.gcap + n] = alpha * sv + xv;
                const float rv = -alpha * q + (pre ? pre_r[a] : p.g_r[a * p.gcap + n]);
                p.g_r[a * p.gcap + n] = rv;
                sum += (double)rv * ((double)rv * m);
                mx += (double)rv * (double)rv;
            }
        }
        wave_reduce2(sum, mx);
        if (ONE) {
            wsum += sum;
            wmx = fmax(wmx, mx);
        } else if (lane == 0) {
            p.bsum[b] = sum;
            p.bmax[b] = mx;
        }
    }
    if (ONE) wg_partials_store(p.bsum, p.bmax, wsum, wmx);
}

// Batches: alpha = rho_k / s.q for every scene still iterating (one
// work-group per scene over the scene's block partials)
// The CG's done poll: Counters::overflow .. cg_iters (five ints) into the
// host's pinned poll slot.
__global__ __launch_bounds__(64) void k_poll(KP p, int* dst) {
    const int t = (int)threadIdx.x;
    if (t < 5) dst[t] = (&p.ctr->overflow)[t];
}

__global__ __launch_bounds__(WG) void k_cg_reduce_sq(KP p, int k) {
    const int si = blockIdx.x;
    CGState& cg = p.cg[si];
    if (cg.done) return;
    const int b0 = p.sst[si].block_begin, b1 = p.sst[si].block_end;
    double sq, unused;
    reduce_range(p.bsq, nullptr, b0, b1 - b0, sq, unused);
    if (threadIdx.x == 0) {
        cg.sq = sq;
        cg.alpha = sq != 0.0 ? cg.rho_hist[k & 1] / sq : (double)FLT_MAX;
        cg.multiplies += 1;
    }
}

// Batches: convergence test of iteration k and beta per scene
__global__ __launch_bounds__(WG) void k_cg_reduce_rr(KP p, int k) {
    const int si = blockIdx.x;
    if (p.cg[si].done) return;
    const int b0 = p.sst[si].block_begin, b1 = p.sst[si].block_end;
    const CGTestIn tin = cg_test_in(p, si, k);
    double rr, mx;
    reduce_range(p.bsum, p.bmax, b0, b1 - b0, rr, mx);
    float beta;
    (void)cg_test(p, si, k, rr, mx, threadIdx.x == 0, beta, tin);
}

// ============================================================================
// G2P (MPM_Example.cpp:631-660): PIC/FLIP velocity, F and X update, removal
// ============================================================================
template <int D>
__global__ __launch_bounds__(BWG) void k_g2p(KP p) {
    using Dm = Dims<D>;
    constexpr int LTN = Dm::LTN, NBR = Dm::NBR, DD = D * D;
    __shared__ float vt[2 * D][LTN];
    __shared__ int nbr[NBR];
    __shared__ unsigned wbest[NW];
    __shared__ int wrem[NW];
    const int nblocks = p.ctr->nblocks;
    const XcdRange xr = xcd_blocks(nblocks);
    for (int kb = xr.first; kb < xr.end; kb += xr.step) {
        const int b = p.block_order[kb];  // heavy blocks first within the XCD's range
        const SceneParams& sc = p.scn[p.block_scene[b]];
        if (!sc.active) continue;
        if (block_empty(p, b)) {  // no particles: empty partials for the scene's reduction
            if (threadIdx.x == 0) {
                p.bu[b] = 0u;
                p.bu[p.bcap + b] = 0u;
            }
            continue;
        }
        const float dt = sc.dt, flip = sc.flip;
        const float lo = 0.0f + 0.5f * sc.dx, hi = 1.0f - 0.5f * sc.dx;
        unsigned best = 0u;
        int removed = 0;
        if (threadIdx.x < NBR) nbr[threadIdx.x] = p.block_nbr[b * NBR + threadIdx.x];
        int bb[D];
        block_coords<D>(p, b, bb);
        // the first particle's position and velocity load while the tile is gathered
        const int i0 = p.block_pstart[b] + (int)threadIdx.x, i1 = p.block_pstart[b + 1];
        float nX[D], nV[D];
        auto load = [&](int i) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                nX[a] = p.pcur[(PF<D>::X + a) * p.pcap + i];
                nV[a] = p.pcur[(PF<D>::V + a) * p.pcap + i];
            }
        };
        if (i0 < i1) load(i0);
        __syncthreads();
        const float* src[2 * D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            src[a] = p.g_vs + a * p.gcap;
            src[D + a] = p.g_v + a * p.gcap;
        }
        gather_tile<D, 2 * D>(p, nbr, vt, src);
        __syncthreads();
        for (int i = i0; i < i1; i += BWG) {
            float w[D][3], dw[D][3], X[D], V0[D];
            int l[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                X[a] = nX[a];
                V0[a] = nV[a];
            }
            if (i + BWG < i1) load(i + BWG);
            bool in_block = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                int cl;
                bspline_weights(X[a], sc.dx, sc.inv_dx, cl, w[a], dw[a]);
                l[a] = cl + Dm::B - Dm::B * bb[a];
                in_block = in_block && l[a] >= 0 && l[a] < Dm::B;
            }
            if (!in_block) continue;
            float vpic[D], vflip[D], G[DD];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                vpic[a] = 0.0f;
                vflip[a] = V0[a];
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
                        for (int cc = 0; cc < D; ++cc) G[a * D + cc] += vs * g[cc];
                    }
                }
            }
            const int mid = sc.nmat > 1 ? __float_as_int(p.pcur[PF<D>::MAT * p.pcap + i]) : 0;
            if (sc.mats[mid].eos) {
                // EOS: density /= 1 + dt tr(grad v) (MPM_Example.cpp:646); F unchanged
                float tr = 0.0f;
#pragma unroll
                for (int a = 0; a < D; ++a) tr += G[a * D + a];
                p.pcur[PF<D>::J * p.pcap + i] = p.pcur[PF<D>::J * p.pcap + i] / (1.0f + dt * tr);
            } else {
                float F[DD];
#pragma unroll
                for (int k = 0; k < DD; ++k) F[k] = p.pcur[(PF<D>::F + k) * p.pcap + i];
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int cc = 0; cc < D; ++cc) {
                        float s = 0.0f;
#pragma unroll
                        for (int k = 0; k < D; ++k) s += (dt * G[r * D + k]) * F[k * D + cc];
                        p.pcur[(PF<D>::F + r * D + cc) * p.pcap + i] = F[r * D + cc] + s;
                    }
            }
            bool inside = true, finite = true;
            float v2 = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float V = vflip[a] * flip + vpic[a] * (1.0f - flip);
                const float Xn = X[a] + vpic[a] * dt;
                p.pcur[(PF<D>::V + a) * p.pcap + i] = V;
                p.pcur[(PF<D>::X + a) * p.pcap + i] = Xn;
                inside = inside && Xn >= lo && Xn <= hi;
                finite = finite && isfinite(V) && isfinite(Xn);
                v2 += V * V;
            }
            // a NaN/Inf state leaves the domain test false (removal, as in the
            // reference) and flags the scene for the caller
            if (!finite) atomicOr(&p.sst[p.block_scene[b]].nonfinite, 1);
            if (!inside) {
                p.keycur[i] = -1;
                ++removed;
            } else {
                best = max(best, __float_as_uint(v2));
            }
        }
        // per-block partials (plain stores); k_g2p_reduce folds each scene's
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            best = max(best, (unsigned)__shfl_xor((int)best, off));
            removed += __shfl_xor(removed, off);
        }
        if ((threadIdx.x & 63) == 0) {
            wbest[threadIdx.x >> 6] = best;
            wrem[threadIdx.x >> 6] = removed;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned bm = 0u;
            int rm = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                bm = max(bm, wbest[w]);
                rm += wrem[w];
            }
            p.bu[b] = bm;
            p.bu[p.bcap + b] = (unsigned)rm;
        }
        __syncthreads();
    }
}

// One work-group per scene: max speed and removal count of the last G2P, in
// a fixed order (MPM_Example::Max_Particle_Velocity, MPM_Example.cpp:375-387).
// Scenes that did not advance keep their values.
__global__ __launch_bounds__(WG) void k_g2p_reduce(KP p) {
    const int si = blockIdx.x;
    if (!p.scn[si].active) return;
    const int b0 = p.sst[si].block_begin, b1 = p.sst[si].block_end;
    constexpr int U = 8;
    unsigned ba[U], ra[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { ba[u] = 0u; ra[u] = 0u; }
    for (int base = b0 + (int)threadIdx.x; base < b1; base += U * WG) {
        // the chunk's loads all go out before the first use (a block past the
        // range reads the range's last one and is not counted)
        unsigned bv[U], rv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = min(base + u * WG, b1 - 1);
            bv[u] = p.bu[b];
            rv[u] = p.bu[p.bcap + b];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * WG < b1) {
                ba[u] = max(ba[u], bv[u]);
                ra[u] += rv[u];
            }
    }
    unsigned best = 0u, removed = 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) { best = max(best, ba[u]); removed += ra[u]; }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        best = max(best, (unsigned)__shfl_xor((int)best, off));
        removed += (unsigned)__shfl_xor((int)removed, off);
    }
    __shared__ unsigned sb[WG / 64], sr[WG / 64];
    if ((threadIdx.x & 63) == 0) {
        sb[threadIdx.x >> 6] = best;
        sr[threadIdx.x >> 6] = removed;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WG / 64; ++w) {
            best = max(best, sb[w]);
            removed += sr[w];
        }
        p.sst[si].max_v2 = best;
        p.sst[si].n_removed = (int)removed;
    }
}

// Max |V|^2 per scene over the particle buffer (before the first binning):
// one atomic per distinct scene in a wave.
template <int D>
__global__ __launch_bounds__(WG) void k_max_speed(KP p) {
    const int n = p.ctr->n_live;
    const int lane = threadIdx.x & 63;
    for (int base = blockIdx.x * WG + (threadIdx.x & ~63); base < n; base += gridDim.x * WG) {
        const int i = base + lane;
        const bool valid = i < n && p.keycur[i] >= 0;
        unsigned v2b = 0u;
        int si = -1;
        if (valid) {
            si = p.scur[i];
            float v2 = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float v = p.pcur[(PF<D>::V + a) * p.pcap + i];
                v2 += v * v;
            }
            v2b = __float_as_uint(v2);
        }
        unsigned long long todo = __ballot(valid);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int ls = __shfl(si, leader);
            const unsigned long long grp = __ballot(valid && si == ls);
            unsigned m = (valid && si == ls) ? v2b : 0u;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
            if (lane == leader && m) atomicMax(&p.sst[ls].max_v2, m);
            todo &= ~grp;
        }
    }
}

// dt and activity of scenes [first, first + n) for the next substep
__global__ __launch_bounds__(WG) void k_set_step(KP p, StepArgs a) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= a.n) return;
    SceneParams& sc = p.scn[a.first + i];
    const float dt = a.dt[i];
    sc.dt = dt > 0.0f ? dt : 0.0f;
    sc.active = (a.force_active || dt > 0.0f) ? 1 : 0;
    if (a.clear_waiting) sc.waiting = 0;
}

// Process_Waiting_Particles ran for every waiting scene: clear the flags.
// After a binning over the block capacity no block was set up and the
// waiting particles have no volume yet: they keep waiting.
__global__ __launch_bounds__(WG) void k_clear_waiting(KP p) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i < p.nscenes && !p.ctr->overflow) p.scn[i].waiting = 0;
}

// Removes every particle of scene `scene` (key -1: the next binning drops it).
__global__ __launch_bounds__(WG) void k_remove_scene(KP p, int scene) {
    const int n = p.ctr->n_live;
    for (int i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG)
        if (p.scur[i] == scene) p.keycur[i] = -1;
}

// The same from a device array of nscenes dts (replayed graphs: the host
// copies the substep's dts into `dt` before each replay).
__global__ __launch_bounds__(WG) void k_set_step_dev(KP p, const float* dt) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= p.nscenes) return;
    SceneParams& sc = p.scn[i];
    const float d = dt[i];
    sc.dt = d > 0.0f ? d : 0.0f;
    sc.active = d > 0.0f ? 1 : 0;
}

// The dt of the next substep of every scene in the device frame loop: the
// reference's rule (MPM_Driver.cpp:55-80) on the scene's max |V| from its last
// G2P (MPM_Example::Max_Particle_Velocity), in the same float operations as
// the host loop, so both take the same dts: dt = clamp(cfl dx / max(vmax,
// 1e-2), min_dt, max_dt); the last substep of the frame takes what remains
// when that is under 1.001 dt, and the one before half of it when it is
// under 2 dt.  A scene at its frame end gets dt 0 (it idles).  One work-group.
__global__ __launch_bounds__(WG) void k_frame_dt(KP p, FrameState* f, float* dt_out) {
#pragma clang fp contract(off)
    int done = 0;
    for (int s = threadIdx.x; s < p.nscenes; s += WG) {
        FrameState st = f[s];
        float d = 0.0f;
        if (!st.done) {
            const float v2 = __uint_as_float(p.sst[s].max_v2);
            const float vmax = (float)sqrt((double)v2);  // the correctly rounded float sqrt
            const float vfloor = vmax < 1e-2f ? 1e-2f : vmax;
            const float cfl_dt = st.cfl * st.dx / vfloor;
            const float capped = cfl_dt < st.max_dt ? cfl_dt : st.max_dt;
            d = st.min_dt < capped ? capped : st.min_dt;
            bool last = false;
            const float rem = st.target - st.time;
            if (rem < d * 1.001f) {
                d = rem;
                last = true;
            } else if (rem < 2.0f * d) {
                d = rem * 0.5f;
            }
            if (!(d > 0.0f)) {
                st.done = 1;
                d = 0.0f;
            } else {
                st.time += d;
                st.steps += 1;
                st.last_dt = d;
                if (last) st.done = 1;
            }
            f[s] = st;
        }
        dt_out[s] = d;
        done += st.done ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) done += __shfl_xor(done, off);
    __shared__ int wsum[WG / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = done;
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
#pragma unroll
        for (int w = 0; w < WG / 64; ++w) total += wsum[w];
        p.ctr->frame_ndone = total;
    }
}

// ============================================================================
// Allocation-free substep: the scene table, the counters and the binning's
// slot and cell counters of the static scratch start over at every call (the
// scratch keeps the last call's contents), and the caller's particle-major
// arrays become binning's field-major input; afterwards every binned
// particle goes back to the caller's arrays at its original index.
// ============================================================================
template <int D>
__global__ __launch_bounds__(WG) void k_noalloc_load(KP p, SceneParams sp, NoallocIn a) {
    const int gid = blockIdx.x * WG + threadIdx.x, stride = gridDim.x * WG;
    if (gid == 0) {
        p.scn[0] = sp;
        Counters c{};
        c.n_live = a.n;
        c.n_prev = a.n;
        c.slot_hw = p.nslots;
        *p.ctr = c;
        p.sst[0] = SceneState{};
        a.d_dt[0] = a.dt;
    }
    for (int s = gid; s < p.nslots; s += stride) {
        p.slot_count[s] = 0;
        p.slot_touched[s] = 0;
        a.slot_scene[s] = 0;
    }
    for (int64_t k = gid; k < (int64_t)p.nslots * 64; k += stride) p.cell_count[k] = 0;
    const Material& mat = sp.mats[0];
    for (int i = gid; i < a.n; i += stride) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            p.pcur[(PF<D>::X + c) * p.pcap + i] = a.x[(int64_t)i * D + c];
            p.pcur[(PF<D>::V + c) * p.pcap + i] = a.v[(int64_t)i * D + c];
        }
#pragma unroll
        for (int k = 0; k < D * D; ++k)
            p.pcur[(PF<D>::F + k) * p.pcap + i] = a.F ? a.F[(int64_t)i * D * D + k] : (k % (D + 1) == 0 ? 1.0f : 0.0f);
        p.pcur[PF<D>::J * p.pcap + i] = a.Jp ? a.Jp[i] : (mat.eos ? mat.rho0 : 1.0f);
        p.pcur[PF<D>::M * p.pcap + i] = a.mass[i];
        p.pcur[PF<D>::VOL * p.pcap + i] = a.volume[i];
        p.pcur[PF<D>::MAT * p.pcap + i] = __int_as_float(0);
        p.idcur[i] = i;
        p.keycur[i] = 0;
        p.scur[i] = 0;
        a.valid[i] = 0;
    }
}

template <int D>
__global__ __launch_bounds__(WG) void k_noalloc_store(KP p, NoallocOut o) {
    const int n = p.ctr->n_live;  // binned particles (dropped ones keep valid = 0)
    for (int j = blockIdx.x * WG + threadIdx.x; j < n; j += gridDim.x * WG) {
        const int64_t i = p.idcur[j];
        o.valid[i] = p.keycur[j] >= 0 ? 1 : 0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            o.x[i * D + c] = p.pcur[(PF<D>::X + c) * p.pcap + j];
            o.v[i * D + c] = p.pcur[(PF<D>::V + c) * p.pcap + j];
        }
        if (o.F) {
#pragma unroll
            for (int k = 0; k < D * D; ++k) o.F[i * D * D + k] = p.pcur[(PF<D>::F + k) * p.pcap + j];
        }
        if (o.Jp) o.Jp[i] = p.pcur[PF<D>::J * p.pcap + j];
        o.volume[i] = p.pcur[PF<D>::VOL * p.pcap + j];
    }
}

// ============================================================================
// launchers
// ============================================================================
template <int D> void launch_noalloc_load(const KP& p, hipStream_t s, int grid, const SceneParams& sp, const NoallocIn& a) {
    hipLaunchKernelGGL(k_noalloc_load<D>, dim3(grid), dim3(WG), 0, s, p, sp, a);
}
template <int D> void launch_noalloc_store(const KP& p, hipStream_t s, int grid, const NoallocOut& o) {
    hipLaunchKernelGGL(k_noalloc_store<D>, dim3(grid), dim3(WG), 0, s, p, o);
}
template void launch_noalloc_load<2>(const KP&, hipStream_t, int, const SceneParams&, const NoallocIn&);
template void launch_noalloc_load<3>(const KP&, hipStream_t, int, const SceneParams&, const NoallocIn&);
template void launch_noalloc_store<2>(const KP&, hipStream_t, int, const NoallocOut&);
template void launch_noalloc_store<3>(const KP&, hipStream_t, int, const NoallocOut&);
void launch_set_step(const KP& p, hipStream_t s, const StepArgs& a) {
    hipLaunchKernelGGL(k_set_step, dim3((a.n + WG - 1) / WG), dim3(WG), 0, s, p, a);
}
void launch_set_step_dev(const KP& p, hipStream_t s, const float* dt) {
    hipLaunchKernelGGL(k_set_step_dev, dim3((p.nscenes + WG - 1) / WG), dim3(WG), 0, s, p, dt);
}
void launch_frame_dt(const KP& p, hipStream_t s, FrameState* f, float* dt) {
    hipLaunchKernelGGL(k_frame_dt, dim3(1), dim3(WG), 0, s, p, f, dt);
}
void launch_clear_waiting(const KP& p, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_waiting, dim3((p.nscenes + WG - 1) / WG), dim3(WG), 0, s, p);
}
void launch_remove_scene(const KP& p, hipStream_t s, int grid, int scene) {
    hipLaunchKernelGGL(k_remove_scene, dim3(grid), dim3(WG), 0, s, p, scene);
}
template <int D> void launch_bin(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_bin<D>, dim3(grid), dim3(WG), 0, s, p); }

// Streaming copy of n4 float4 (the measured bandwidth roof): one float4
// per thread over a grid that covers the buffer (on MI355X this form beat
// grid-stride loops with four loads in flight, 6.2 against 4.3-4.9 TB/s).
__global__ __launch_bounds__(WG) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}
void launch_copy(const float4* src, float4* dst, int64_t n4, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n4 + WG - 1) / WG)), dim3(WG), 0, s, src, dst, n4);
}
void launch_scan(const KP& p, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(p.nparts), dim3(WG), 0, s, p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(WG), 0, s, p);
    hipLaunchKernelGGL(k_scan_apply, dim3(p.nparts), dim3(WG), 0, s, p);
}
template <int D> void launch_block_setup(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_block_setup<D>, dim3(grid), dim3(WG), 0, s, p); }
void launch_setup_order(const KP& p, hipStream_t s, int cell_groups) {
    const int grid = ORDER_GROUPS + (p.nscenes + WG - 1) / WG + cell_groups;
    hipLaunchKernelGGL(k_setup_order, dim3(grid), dim3(WG), 0, s, p);
}
void launch_poll(const KP& p, hipStream_t s, int* dst) { hipLaunchKernelGGL(k_poll, dim3(1), dim3(64), 0, s, p, dst); }
template <int D> void launch_reorder(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_reorder<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_p2g(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_p2g<D>, dim3(grid), dim3(BWG), 0, s, p); }
template <int D> void launch_normalize(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_normalize<D>, dim3(grid), dim3(WG), 0, s, p); }
template <int D> void launch_init_volume(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_init_volume<D>, dim3(grid), dim3(BWG), 0, s, p); }
template <int D> void launch_particle_force(const KP& p, hipStream_t s, int grid, bool constitutive, bool force) {
    if (constitutive && force) hipLaunchKernelGGL((k_particle_force<D, true, true>), dim3(grid), dim3(BWG), 0, s, p);
    else if (constitutive) hipLaunchKernelGGL((k_particle_force<D, true, false>), dim3(grid), dim3(BWG), 0, s, p);
    else if (force) hipLaunchKernelGGL((k_particle_force<D, false, true>), dim3(grid), dim3(BWG), 0, s, p);
}
template <int D> void launch_grid_update(const KP& p, hipStream_t s, int grid, bool force, bool collision, bool cg_init, bool write_f) {
    if (force && collision && !cg_init && !write_f) hipLaunchKernelGGL((k_grid_update<D, true, true, false, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (force && !collision && !cg_init && write_f) hipLaunchKernelGGL((k_grid_update<D, true, false, false, true>), dim3(grid), dim3(WG), 0, s, p);
    else if (force && !collision && !cg_init) hipLaunchKernelGGL((k_grid_update<D, true, false, false, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (!force && collision && !cg_init) hipLaunchKernelGGL((k_grid_update<D, false, true, false, false>), dim3(grid), dim3(WG), 0, s, p);
    else if (!force && !collision && cg_init) hipLaunchKernelGGL((k_grid_update<D, false, false, true, false>), dim3(grid), dim3(WG), 0, s, p);
    else hipLaunchKernelGGL((k_grid_update<D, true, true, true, false>), dim3(grid), dim3(WG), 0, s, p);
}
template <int D> void launch_cg_force(const KP& p, hipStream_t s, int grid, bool restart, int k, bool one) {
    if (restart) hipLaunchKernelGGL((k_cg_force<D, true, false>), dim3(grid), dim3(BWG), 0, s, p, 0);
    else if (one) hipLaunchKernelGGL((k_cg_force<D, false, true>), dim3(grid), dim3(BWG), 0, s, p, k);
    else hipLaunchKernelGGL((k_cg_force<D, false, false>), dim3(grid), dim3(BWG), 0, s, p, k);
}
template <int D> void launch_cg_force_cell(const KP& p, hipStream_t s, int grid, bool restart, int k, bool one) {
    if (restart) hipLaunchKernelGGL((k_cg_force_cell<D, true, false>), dim3(grid), dim3(WG), 0, s, p, 0);
    else if (one) hipLaunchKernelGGL((k_cg_force_cell<D, false, true>), dim3(grid), dim3(WG), 0, s, p, k);
    else hipLaunchKernelGGL((k_cg_force_cell<D, false, false>), dim3(grid), dim3(WG), 0, s, p, k);
}
template <int D> void launch_cg_restart_node(const KP& p, hipStream_t s, int grid, bool one) {
    if (one) hipLaunchKernelGGL((k_cg_restart_node<D, true>), dim3(grid), dim3(WG), 0, s, p);
    else hipLaunchKernelGGL((k_cg_restart_node<D, false>), dim3(grid), dim3(WG), 0, s, p);
}
template <int D> void launch_cg_update_s(const KP& p, hipStream_t s, int grid, int k) {
    hipLaunchKernelGGL(k_cg_update_s<D>, dim3(grid), dim3(WG), 0, s, p, k, grid);
}
template <int D> void launch_cg_node1(const KP& p, hipStream_t s, int grid, int k, bool one) {
    if (one) hipLaunchKernelGGL((k_cg_node1<D, true>), dim3(grid), dim3(WG), 0, s, p, k);
    else hipLaunchKernelGGL((k_cg_node1<D, false>), dim3(grid), dim3(WG), 0, s, p, k);
}
template <int D> void launch_cg_node2(const KP& p, hipStream_t s, int grid, int k, bool one) {
    if (one) hipLaunchKernelGGL((k_cg_node2<D, true>), dim3(grid), dim3(WG), 0, s, p, k, grid);
    else hipLaunchKernelGGL((k_cg_node2<D, false>), dim3(grid), dim3(WG), 0, s, p, k, grid);
}
void launch_cg_reduce_sq(const KP& p, hipStream_t s, int k) { hipLaunchKernelGGL(k_cg_reduce_sq, dim3(p.nscenes), dim3(WG), 0, s, p, k); }
void launch_cg_reduce_rr(const KP& p, hipStream_t s, int k) { hipLaunchKernelGGL(k_cg_reduce_rr, dim3(p.nscenes), dim3(WG), 0, s, p, k); }
template <int D> void launch_g2p(const KP& p, hipStream_t s, int grid) {
    hipLaunchKernelGGL(k_g2p<D>, dim3(grid), dim3(BWG), 0, s, p);
    hipLaunchKernelGGL(k_g2p_reduce, dim3(p.nscenes), dim3(WG), 0, s, p);
}
template <int D> void launch_max_speed(const KP& p, hipStream_t s, int grid) { hipLaunchKernelGGL(k_max_speed<D>, dim3(grid), dim3(WG), 0, s, p); }

#define MPM_INSTANTIATE(D)                                                                    \
    template void launch_bin<D>(const KP&, hipStream_t, int);                                 \
    template void launch_block_setup<D>(const KP&, hipStream_t, int);                         \
    template void launch_reorder<D>(const KP&, hipStream_t, int);                             \
    template void launch_p2g<D>(const KP&, hipStream_t, int);                                 \
    template void launch_normalize<D>(const KP&, hipStream_t, int);                           \
    template void launch_init_volume<D>(const KP&, hipStream_t, int);                         \
    template void launch_particle_force<D>(const KP&, hipStream_t, int, bool, bool);          \
    template void launch_grid_update<D>(const KP&, hipStream_t, int, bool, bool, bool, bool); \
    template void launch_cg_force<D>(const KP&, hipStream_t, int, bool, int, bool);           \
    template void launch_cg_force_cell<D>(const KP&, hipStream_t, int, bool, int, bool);      \
    template void launch_cg_restart_node<D>(const KP&, hipStream_t, int, bool);               \
    template void launch_cg_update_s<D>(const KP&, hipStream_t, int, int);                    \
    template void launch_cg_node1<D>(const KP&, hipStream_t, int, int, bool);                 \
    template void launch_cg_node2<D>(const KP&, hipStream_t, int, int, bool);                 \
    template void launch_g2p<D>(const KP&, hipStream_t, int);                                 \
    template void launch_max_speed<D>(const KP&, hipStream_t, int);
MPM_INSTANTIATE(2)
MPM_INSTANTIATE(3)

}  // namespace mpm
