// Penalty contact between deformable surfaces on the GPU: solver/Contact.java:132-763.
//
// Included by neofiz_hip.hip after the Dev and Ctl definitions.
//
// Each assembly that gathers forces runs four launches after the gather and before the
// nodal update (Contact.begin / search / accumulate, :424-462):
//
//   (c_rebuild, on extra workgroups of the gather launch) only when an element was forgotten
//                 since the last rebuild (adopt, :290-322): live flags and the compact list of
//                 every potential segment.
//   k_c_nodes_geometry
//                 the rest of that rebuild (live flags and compact list of surface nodes,
//                 lastMaster cleared) beside geometry: segment lengths and outward normals
//                 (:465-487), the grid's bounding box and longest segment (exact min/max on
//                 order-preserving bit keys), and the clearing of the hash heads and
//                 per-segment pair lists.
//   k_c_grid      node normals in the reference's order (:488-499) and every segment end
//                 filed under its grid cell (:515-551).
//   k_c_search    per surface node, the deepest penetrated segment (:570-658) and the
//                 accumulate bookkeeping (:446-462); the pair's forces are evaluated
//                 (resolve, :660-762) and the slave is linked to its master segment.
//   k_c_apply     per node, the pair forces that touch it added to the gathered force in
//                 the reference's order: by slave node, then normal before tangent.
//
// Identity instead of order. The reference walks a surface list built in element order and
// a bucket grid built in segment order, and several results depend on that order: the first
// of two equally deep candidates wins, and each node's force takes its contact terms in
// slave order. Here every segment is named by its potential index p = 4 e + k, which sorts
// exactly as the reference's surface index does, and every surface node by its node index,
// which sorts as the reference's surface-node index does. The grid is a hash of the
// reference's own cells (same origin, cell size and column count), whose buckets may hold
// entries of other cells; only entries of the nine reference cells are considered, and the
// first-visit position of a candidate is reconstructed as (cell rank in the 3 x 3 walk,
// entry index 2p + end). The winner is then the deepest candidate with the smallest such
// position, which is the candidate the reference's sequential scan keeps. So every pair,
// every force and every escape count is the reference's, bit for bit, while the lists the
// kernels walk are built with atomics in any order.
//
// The per-step audits (pairs, stored energy, maximum penetration) and the running friction
// and damper dissipation are reduced per launch in a fixed tree (node-index blocks), not in
// the reference's serial pair order, so the two energies agree with the reference to
// rounding rather than bitwise; the counts and the maximum are exact.
#pragma once

constexpr double C_MAX_DEPTH = 0.5;         // Contact.java:153
constexpr double C_STICK_RELAXATION = 0.5;  // :171
constexpr double C_FACING = 0.25;           // :195
constexpr int CPAIR = 10;                   // doubles per resolved pair record
constexpr int C_MAX_TERMS = 32;             // pair terms one node may receive in one step
constexpr int CONTACT_GRID = 1024;          // blocks of the grid-stride contact launches

// A pair record: what resolve() computed, read back by the nodes it pushes on.
enum { PR_NORMAL = 0, PR_TANGENT, PR_WA, PR_WB, PR_NR, PR_NZ, PR_ENERGY, PR_DAMPER, PR_FRICTION, PR_DEPTH };

__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long b = (unsigned long long)nfm::bits_of(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__device__ __forceinline__ double from_order_key(unsigned long long k) {
    return nfm::from_bits((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k));
}

// Java's (int) cast of a double: truncation toward zero, saturating, NaN to 0.
__device__ __forceinline__ int jint(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int)x;
}
__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : v >= n ? n - 1 : v; }

__device__ __forceinline__ void seg_ends(const Dev& d, int p, int& a, int& b) {
    const int4 c = d.conn[p >> 2];
    const int k = p & 3;
    a = k == 0 ? c.x : k == 1 ? c.y : k == 2 ? c.z : c.w;
    b = k == 0 ? c.y : k == 1 ? c.z : k == 2 ? c.w : c.x;
}

// Current position, cr = r + ur, cz = z + uz (:467-470).
__device__ __forceinline__ double2 cpos(const Dev& d, int n) {
    const double2 x = d.x0[n], u = d.u[n];
    return make_double2(x.x + u.x, x.y + u.y);
}

// p is on the surface of the elements the contact still knows: its element is live and no
// other live element uses the same edge (QuadMesh.surface, mesh/QuadMesh.java:188-226).
// An edge of a conforming quad mesh has at most one partner, read from cPartner: two
// dependent loads. An edge with several partners walks the CSR.
__device__ __forceinline__ bool seg_live(const Dev& d, int p) {
    bool live = d.cDead[p >> 2] == 0;
    const int partner = d.cPartner[p];
    if (partner >= 0) {
        live &= d.cDead[partner >> 2] != 0;
    } else if (partner == -2) {
        for (int k = d.cShareStart[p]; k < d.cShareStart[p + 1]; k++) live &= d.cDead[d.cShareOf[k] >> 2] != 0;
    }
    return live;
}

// Node n's first eight segments (all of them on a quad mesh) from the fixed-width copy, -1
// where it has fewer.
__device__ __forceinline__ void node_segs8(const Dev& d, int n, int (&seg)[8]) {
    const int4 a = d.cNodeSeg8[2 * n], b = d.cNodeSeg8[2 * n + 1];
    seg[0] = a.x; seg[1] = a.y; seg[2] = a.z; seg[3] = a.w;
    seg[4] = b.x; seg[5] = b.y; seg[6] = b.z; seg[7] = b.w;
}

// Whether node n touches a live segment: the first eight segments tested together, any
// further ones through the CSR.
__device__ __forceinline__ bool node_on_surface(const Dev& d, int n) {
    int seg[8];
    node_segs8(d, n, seg);
    bool on = false;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (seg[k] >= 0) on |= d.cActive[seg[k]] != 0;
    const int k0 = d.cNodeSegStart[n], k1 = d.cNodeSegStart[n + 1];
    for (int k = k0 + 8; k < k1; k++) on |= d.cActive[d.cNodeSegOf[k]] != 0;
    return on;
}

struct CGrid {
    double r0, z0, cell;
    int cols, rows;
    unsigned mask;  // hash table size - 1
};

__device__ __forceinline__ unsigned c_hash_size(int nAct) {
    unsigned h = 64;
    while (h < 4u * (unsigned)nAct) h <<= 1;
    return h;
}

// Contact.grid :515-533, from the reduced bounding box and longest segment.
__device__ __forceinline__ CGrid c_grid(const Dev& d) {
    const Ctl* c = d.ctl;
    CGrid g;
    const double rlo = from_order_key(c->cBox[0]), zlo = from_order_key(c->cBox[1]);
    const double rhi = -from_order_key(c->cBox[2]), zhi = -from_order_key(c->cBox[3]);
    const double longest = nfm::from_bits((long long)c->cLongest);
    g.cell = longest > 0.0 ? 1.5 * longest : 1.0;
    g.r0 = rlo;
    g.z0 = zlo;
    const int ci = (int)((unsigned)jint((rhi - rlo) / g.cell) + 1u);
    const int cj = (int)((unsigned)jint((zhi - zlo) / g.cell) + 1u);
    g.cols = ci > 1 ? ci : 1;
    g.rows = cj > 1 ? cj : 1;
    g.mask = c_hash_size(c->cNAct) - 1u;
    return g;
}

__device__ __forceinline__ unsigned c_bucket(int i, int j, unsigned mask) {
    unsigned h = (unsigned)i * 0x9E3779B1u ^ (unsigned)j * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h & mask;
}

// Capture-grid hash table size: 4 buckets per live segment (each bucket keeps C_SUB lists,
// so 4 C_SUB heads), within what was allocated. The heads are cleared every assembly.
__device__ __forceinline__ unsigned c_capture_size(const Dev& d, int nAct) {
    unsigned h = 64;
    while (h < 4u * (unsigned)nAct && h < d.cFHashMax) h <<= 1;
    return h;
}

// The capture-grid cell holding coordinate x (cells anchored at the grid origin).
__device__ __forceinline__ int c_fine(double x, double origin, double h) {
    const double f = __builtin_floor((x - origin) / h);
    return f < -1.0e9 ? -1000000000 : (f > 1.0e9 ? 1000000000 : (int)f);
}

constexpr int C_MAX_CAPTURE_CELLS = 1 << 16;  // per segment; beyond it the step falls back
// Each capture bucket keeps C_SUB lists (an entry goes to list `entry mod C_SUB`), and the
// search walks them with C_SUB lanes per surface node, so no lane follows a long chain.
constexpr int C_SUB = 16;

__device__ __forceinline__ bool c_skip(const Dev& d, int q, int checkGo) {
    return checkGo && !d.ctl->s[q].go;
}

// ------------------------------------------------------------------ launches

// Exclusive prefix of v over the workgroup in thread order, and the workgroup's total.
__device__ __forceinline__ int block_exclusive(int v, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < BLOCK / 64; w++) {
        if (w < wave) before += lds[w];
        total += lds[w];
    }
    __syncthreads();
    return before + x - v;
}

// The rebuild launches give each thread a few items (strided by the workgroup, so loads and
// flag stores stay coalesced) and reserve a workgroup's appends with one atomic: a counter
// shared by the whole grid serialises its atomics, so they are kept to a few hundred. The
// segment pass takes more items per thread (more workgroups would mean more atomics on its
// counter), the node pass fewer (its dependent loads want more waves in flight).
constexpr int C_CHUNK_SEG = 8;
constexpr int C_CHUNK_NODE = 2;

// Appends this thread's items (bit j of mask: item first + j * BLOCK) to list[0, cap) through
// counter, one reservation per workgroup.
template <int CHUNK>
__device__ __forceinline__ void block_append(unsigned mask, int first, int* counter, int* list, int cap) {
    __shared__ int lds[BLOCK / 64];
    __shared__ int ldsBase;
    int total;
    int at = block_exclusive(__popc(mask), lds, total);
    if (threadIdx.x == 0) ldsBase = total ? atomicAdd(counter, total) : 0;
    __syncthreads();
    at += ldsBase;
    for (int j = 0; j < CHUNK; j++)
        if ((mask >> j) & 1u) {
            if (at >= 0 && at < cap) list[at] = first + j * BLOCK;
            at++;
        }
}

// adopt() :290-322 when the contact has forgotten elements since the last rebuild (workgroup
// `block` of the segment part). The counters were zeroed by whoever raised the stale flag.
// Every assembly: the grid's reductions are reset here, ahead of the geometry part of
// k_c_nodes_geometry. Inside a step this runs on extra workgroups of the gather launch,
// since it reads only the elements' forgotten flags.
__device__ void c_rebuild(const Dev& d, int block) {
    Ctl* c = d.ctl;
    if (block == 0 && threadIdx.x == 0) {
        const unsigned long long hi = order_key(NFD_MAX_VALUE);
        c->cBox[0] = hi;
        c->cBox[1] = hi;
        c->cBox[2] = hi;  // keys of -rhi and -zhi: both start at -(-MAX_VALUE)
        c->cBox[3] = hi;
        c->cLongest = 0ULL;
        c->cFCount = 0;
        c->cFOverflow = 0;
    }
    if (!c->surfaceStale) return;
    const int E4 = 4 * d.E;
    const int first = block * BLOCK * C_CHUNK_SEG + threadIdx.x;
    unsigned mask = 0;
#pragma unroll
    for (int j = 0; j < C_CHUNK_SEG; j++) {
        const int p = first + j * BLOCK;
        if (p < E4) {
            const bool live = seg_live(d, p);
            d.cActive[p] = live ? 1 : 0;
            mask |= (live ? 1u : 0u) << j;
        }
    }
    block_append<C_CHUNK_SEG>(mask, first, &c->cNAct, d.cAct, E4);
}

__global__ void __launch_bounds__(BLOCK) k_c_rebuild(Dev d, int q, int checkGo) {
    if (c_skip(d, q, checkGo)) return;
    c_rebuild(d, blockIdx.x);
}

// The second half of the rebuild: surface nodes from the live flags just written (workgroup
// `block` of the node part).
__device__ void c_rebuild_nodes(const Dev& d, int block) {
    Ctl* c = d.ctl;
    const int first = block * BLOCK * C_CHUNK_NODE + threadIdx.x;
    unsigned mask = 0;
#pragma unroll
    for (int j = 0; j < C_CHUNK_NODE; j++) {
        const int n = first + j * BLOCK;
        if (n < d.N) {
            const bool on = node_on_surface(d, n);
            d.cSurfFlag[n] = on ? 1 : 0;
            d.cPairOk[n] = 0;
            d.cLast[n] = -1;
            mask |= (on ? 1u : 0u) << j;
        }
    }
    block_append<C_CHUNK_NODE>(mask, first, &c->cNSurf, d.cSurf, d.N);
}

// geometry() :465-487 per live segment; the grid's box and longest segment; clears
// (workgroup `block` of `nblocks` in the geometry part).
__device__ void c_geometry(const Dev& d, int block, int nblocks) {
    Ctl* c = d.ctl;
    const int nAct = c->cNAct;
    const int stride = nblocks * BLOCK;
    const int t0 = block * BLOCK + threadIdx.x;
    unsigned long long box[4] = {~0ULL, ~0ULL, ~0ULL, ~0ULL};
    unsigned long long longest = 0ULL;
    for (int t = t0; t < nAct; t += stride) {
        const int p = d.cAct[t];
        int a, b;
        seg_ends(d, p, a, b);
        const double2 xa = cpos(d, a), xb = cpos(d, b);
        const double dr = xb.x - xa.x, dz = xb.y - xa.y;
        const double len = __builtin_sqrt(dr * dr + dz * dz);
        d.cLen[p] = len;
        d.cNrm[p] = len > 0.0 ? make_double2(dz / len, -dr / len) : make_double2(0.0, 0.0);
        d.cSegX[p] = make_double4(xa.x, xa.y, xb.x, xb.y);
        d.cPairHead[p] = -1;
        // Both ends of every live segment are exactly the surface nodes. NaN never wins a
        // comparison in the reference's loops, so it is left out here.
        const double xs[4] = {xa.x, xa.y, xb.x, xb.y};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double v = xs[k];
            if (v != v) continue;
            const unsigned long long lo = order_key(v), hi = order_key(-v);
            const int axis = k & 1;
            box[axis] = lo < box[axis] ? lo : box[axis];
            box[2 + axis] = hi < box[2 + axis] ? hi : box[2 + axis];
        }
        if (len == len) {
            const unsigned long long lb = (unsigned long long)nfm::bits_of(len);
            longest = lb > longest ? lb : longest;
        }
    }
    // Wave shuffles, then the workgroup's waves through LDS: one atomic per workgroup and
    // key (the keys share a line, so the grid's atomics on them serialise).
    __shared__ unsigned long long ldsKey[BLOCK / 64][5];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        unsigned long long v = box[k];
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(v, off, 64);
            v = o < v ? o : v;
        }
        if ((threadIdx.x & 63) == 0) ldsKey[wave][k] = v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(longest, off, 64);
        longest = o > longest ? o : longest;
    }
    if ((threadIdx.x & 63) == 0) ldsKey[wave][4] = longest;
    __syncthreads();
    if (threadIdx.x < 5) {
        const int k = threadIdx.x;
        unsigned long long v = ldsKey[0][k];
        for (int w = 1; w < BLOCK / 64; w++) {
            const unsigned long long o = ldsKey[w][k];
            v = k < 4 ? (o < v ? o : v) : (o > v ? o : v);
        }
        if (k < 4 && v != ~0ULL) atomicMin(&c->cBox[k], v);
        if (k == 4 && v) atomicMax(&c->cLongest, v);
    }
    const unsigned H = c_hash_size(nAct);
    for (unsigned h = t0; h < H; h += stride) d.cHead[h] = -1;
    const unsigned HF = c_capture_size(d, nAct) * C_SUB;
    for (unsigned h = t0; h < HF; h += stride) d.cFHead[h] = -1;
}

// The rest of the rebuild (surface nodes, only after an erosion) and geometry() side by side:
// they read different outputs of k_c_rebuild and write different arrays, so the first
// nodeBlocks workgroups do the one and the others the other, and their loads overlap.
__global__ void __launch_bounds__(BLOCK) k_c_nodes_geometry(Dev d, int q, int checkGo, int nodeBlocks) {
    if (c_skip(d, q, checkGo)) return;
    if ((int)blockIdx.x < nodeBlocks) {
        if (d.ctl->surfaceStale) c_rebuild_nodes(d, blockIdx.x);
        return;
    }
    c_geometry(d, blockIdx.x - nodeBlocks, gridDim.x - nodeBlocks);
}

// Node normals :488-499 (the node's live segments in surface order) and the grid :534-551.
// The launch is three equal thirds, one per independent part (node normals, the coarse grid,
// the capture grid), so that their dependent loads overlap instead of queueing in each thread.
__global__ void __launch_bounds__(BLOCK) k_c_grid(Dev d, int q, int checkGo) {
    if (c_skip(d, q, checkGo)) return;
    Ctl* c = d.ctl;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->surfaceStale = 0;  // the rebuild has been read
    const int nAct = c->cNAct, nSurf = c->cNSurf;
    const int third = gridDim.x / 3;
    const int part = blockIdx.x / third;
    const int block = blockIdx.x - part * third;
    const int stride = third * BLOCK;
    const int t0 = block * BLOCK + threadIdx.x;
    if (part == 0) for (int t = t0; t < nSurf; t += stride) {
        const int n = d.cSurf[t];
        double sr = 0.0, sz = 0.0;
        // the first eight segments' flags and normals loaded side by side, summed in order
        const int k0 = d.cNodeSegStart[n], k1 = d.cNodeSegStart[n + 1];
        int seg8[8];
        double2 nv8[8];
        node_segs8(d, n, seg8);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool live = seg8[k] >= 0 && d.cActive[seg8[k]];
            nv8[k] = live ? d.cNrm[seg8[k]] : make_double2(0.0, 0.0);
            if (!live) seg8[k] = -1;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (seg8[k] >= 0) {
                sr += nv8[k].x;
                sz += nv8[k].y;
            }
        for (int k = k0 + 8; k < k1; k++) {
            const int p = d.cNodeSegOf[k];
            if (!d.cActive[p]) continue;
            const double2 nv = d.cNrm[p];
            sr += nv.x;
            sz += nv.y;
        }
        const double len = __builtin_sqrt(sr * sr + sz * sz);
        d.cPn[n] = len > 0.0 ? make_double2(sr / len, sz / len) : make_double2(0.0, 0.0);
    }
    if (part == 0 || nAct == 0) return;
    const CGrid g = c_grid(d);
    if (part == 1) for (int t = t0; t < 2 * nAct; t += stride) {
        const int p = d.cAct[t >> 1];
        int a, b;
        seg_ends(d, p, a, b);
        const double2 x = cpos(d, (t & 1) ? b : a);
        const int i = clampi(jint((x.x - g.r0) / g.cell), g.cols);
        const int j = clampi(jint((x.y - g.z0) / g.cell), g.rows);
        const int entry = 2 * p + (t & 1);
        d.cEntCell[entry] = make_int2(i, j);
        d.cNext[entry] = atomicExch(&d.cHead[c_bucket(i, j, g.mask)], entry);
    }
    // The capture grid. A segment can be a node's master only if the node lies within
    // its capture rectangle: projection along it in [0, len], depth below it in (0, len / 2]
    // (depth() :639-649 with the MAX_DEPTH cut). That rectangle lies inside the segment's
    // bounding box grown by len / 2 (a hair more, for rounding), and the segment is listed
    // in every cell that box overlaps, so the node's own cell lists every candidate.
    // Entries are reserved once per workgroup round (the order of a cell's list is
    // immaterial: the search ranks candidates by their walk key); every thread of a
    // workgroup runs the same rounds, so the reservation's barriers are uniform.
    if (part != 2) return;
    const unsigned fmask = c_capture_size(d, nAct) - 1u;
    __shared__ int lds[BLOCK / 64];
    __shared__ int ldsBase;
    for (int round = block * BLOCK; round < nAct; round += stride) {
        const int t = round + threadIdx.x;
        int p = -1, count = 0, i0 = 0, i1 = -1, j0 = 0;
        if (t < nAct) {
            p = d.cAct[t];
            const double len = d.cLen[p];
            if (len > 0.0 && len <= NFD_MAX_VALUE) {  // any other segment never wins
                int a, b;
                seg_ends(d, p, a, b);
                const double2 xa = cpos(d, a), xb = cpos(d, b);
                const double m = 0.5 * len * (1.0 + 1.0e-6) + 1.0e-12;
                i0 = c_fine(nfd::jmin(xa.x, xb.x) - m, g.r0, d.cFine);
                i1 = c_fine(nfd::jmax(xa.x, xb.x) + m, g.r0, d.cFine);
                j0 = c_fine(nfd::jmin(xa.y, xb.y) - m, g.z0, d.cFine);
                const int j1 = c_fine(nfd::jmax(xa.y, xb.y) + m, g.z0, d.cFine);
                const long long cells = (long long)(i1 - i0 + 1) * (j1 - j0 + 1);
                if (cells > C_MAX_CAPTURE_CELLS) atomicOr(&c->cFOverflow, 1);
                else count = (int)cells;
            }
        }
        int total;
        const int offset = block_exclusive(count, lds, total);
        if (threadIdx.x == 0) ldsBase = total ? atomicAdd(&c->cFCount, total) : 0;
        __syncthreads();
        const int base = ldsBase;
        __syncthreads();  // ldsBase is rewritten by the next round
        if (count == 0) continue;
        if (base < 0 || (long long)base + offset + count > d.cFCap) {  // full (or the counter wrapped)
            atomicOr(&c->cFOverflow, 1);
            continue;
        }
        // Up to four cells (the usual count) have their head exchanges issued together
        // before the entries are stored; larger boxes continue one cell at a time.
        const int width = i1 - i0 + 1;
        const int at0 = base + offset;
        int next4[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < count) {
                const int at = at0 + k, i = i0 + k % width, j = j0 + k / width;
                next4[k] = atomicExch(&d.cFHead[c_bucket(i, j, fmask) * C_SUB + (at & (C_SUB - 1))], at);
            }
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < count) d.cFEnt[at0 + k] = make_int4(next4[k], p, i0 + k % width, j0 + k / width);
        for (int k = 4, at = at0 + 4; k < count; k++, at++) {
            const int i = i0 + k % width, j = j0 + k / width;
            const int next = atomicExch(&d.cFHead[c_bucket(i, j, fmask) * C_SUB + (at & (C_SUB - 1))], at);
            d.cFEnt[at] = make_int4(next, p, i, j);
        }
    }
}

// excluded() :617-636
__device__ __forceinline__ bool c_excluded(const Dev& d, int n, double2 pn, int p, int a, int b) {
    const double2 ns = d.cNrm[p];
    if (pn.x * ns.x + pn.y * ns.y > -C_FACING) return true;
    if (a == n || b == n) return true;
    for (int k = d.cNodeSegStart[n]; k < d.cNodeSegStart[n + 1]; k++) {
        const int own = d.cNodeSegOf[k];
        if (!d.cActive[own]) continue;
        int pa, pb;
        seg_ends(d, own, pa, pb);
        if (pa == a || pa == b || pb == a || pb == b) return true;
    }
    return false;
}

// depth() :639-649, with the segment's end positions, length and normal given.
__device__ __forceinline__ double c_depth_of(double2 xn, double2 xa, double2 xb, double len, double2 ns) {
    if (len <= 0.0) return 0.0;
    const double tr = (xb.x - xa.x) / len, tz = (xb.y - xa.y) / len;
    const double dr = xn.x - xa.x, dz = xn.y - xa.y;
    const double along = dr * tr + dz * tz;
    if (along < 0.0 || along > len) return 0.0;
    const double gap = dr * ns.x + dz * ns.y;
    return gap < 0.0 ? -gap : 0.0;
}

__device__ __forceinline__ double c_depth_at(const Dev& d, double2 xn, int p, double2 xa, double2 xb) {
    return c_depth_of(xn, xa, xb, d.cLen[p], d.cNrm[p]);
}

__device__ __forceinline__ double c_depth(const Dev& d, double2 xn, int p, int a, int b) {
    return c_depth_at(d, xn, p, cpos(d, a), cpos(d, b));
}

// nearestPenetrated :570-606 and accumulate :446-462 for one surface node, with resolve
// :660-762 evaluated into the node's pair record.
// schedule: the last workgroup makes the element pass's schedule instead (as the node pass's
// extra workgroup does without contact: from the costs of the element launch that read Step
// q, for the one that reads Step q ^ 1), and the others stride without it.
__global__ void __launch_bounds__(BLOCK) k_c_search(Dev d, int q, int checkGo, int schedule) {
    if (schedule && blockIdx.x == gridDim.x - 1) {
        schedule_patches(d, d.patchCost + (size_t)q * d.patches, d.patchOf + (size_t)(q ^ 1) * 4 * d.patchBlocks,
                         d.patchRun ? d.patchRun + (size_t)(q ^ 1) * 4 * d.patchBlocks : nullptr);
        return;
    }
    if (c_skip(d, q, checkGo)) return;
    Ctl* c = d.ctl;
    const int nAct = c->cNAct, nSurf = c->cNSurf;
    if (nAct == 0) return;
    const CGrid g = c_grid(d);
    const unsigned fmask = c_capture_size(d, nAct) - 1u;
    const double dt = c->s[q].dt;
    const int stride = (gridDim.x - schedule) * BLOCK;
    long long escaped = 0;
    // C_SUB consecutive lanes per surface node (stride is a multiple of C_SUB, so a group
    // stays together): lane `sub` walks list `sub` of the node's capture bucket, the group
    // then keeps the best of its candidates, and its first lane carries on alone.
    for (int t = blockIdx.x * BLOCK + threadIdx.x; t < C_SUB * nSurf; t += stride) {
        const int sub = t & (C_SUB - 1);
        const int n = d.cSurf[t / C_SUB];
        const double2 xn = cpos(d, n);
        const double2 pn = d.cPn[n];
        const int last = d.cLast[n];
        const int ci = clampi(jint((xn.x - g.r0) / g.cell), g.cols);
        const int cj = clampi(jint((xn.y - g.z0) / g.cell), g.rows);
        int best = -1;
        double deepest = 0.0;
        unsigned long long bestKey = ~0ULL;
        bool fell = false;
        // One candidate as the reference's loop body treats it (:585-602). The reference
        // tests the exclusions first, then the depth; both are pure, and a candidate is kept
        // only if it passes both, so testing the cheaper depth first selects the same ones.
        // xa, xb: the segment's current end positions (cSegX, or cpos where that is stale).
        auto consider = [&](int p, double2 xa, double2 xb, double len, double2 ns, unsigned long long key) {
            const double dd = c_depth_of(xn, xa, xb, len, ns);
            if (dd <= 0.0) return;
            int a, b;
            seg_ends(d, p, a, b);
            if (c_excluded(d, n, pn, p, a, b)) return;
            if (dd > C_MAX_DEPTH * len) {
                if (p == last) fell = true;
                return;
            }
            if (dd > deepest || (dd == deepest && key < bestKey)) {
                deepest = dd;
                best = p;
                bestKey = key;
            }
        };
        // Where segment p's first entry falls in the reference's walk over the nine cells
        // around (ci, cj): (cell rank, entry index 2p + end), or ~0 if neither end is there.
        // (Each end through its own call, so no array of the two ends reaches scratch memory.)
        auto endKey = [&](double2 x, unsigned long long entry) {
            const int di = clampi(jint((x.x - g.r0) / g.cell), g.cols) - ci;
            const int dj = clampi(jint((x.y - g.z0) / g.cell), g.rows) - cj;
            if (di < -1 || di > 1 || dj < -1 || dj > 1) return ~0ULL;
            return ((unsigned long long)((dj + 1) * 3 + (di + 1)) << 40) | entry;
        };
        auto walkKey = [&](int p, double2 xa, double2 xb) {
            const unsigned long long ka = endKey(xa, (unsigned long long)(2 * p));
            const unsigned long long kb = endKey(xb, (unsigned long long)(2 * p + 1));
            return kb < ka ? kb : ka;
        };
        if (!c->cFOverflow) {
            // The node's capture cell lists every segment that can be its master; each is
            // considered only if the reference's walk reaches it, at its walk position.
            const int fi = c_fine(xn.x, g.r0, d.cFine), fj = c_fine(xn.y, g.z0, d.cFine);
            // The walk keeps the next link's load in flight while a candidate's own data
            // (end positions, length, normal: three loads issued together) is read.
            // (An exhausted link loads entry 0 and ignores it: a select between a load and a
            // constant would put the constant in scratch memory.)
            int x = d.cFHead[c_bucket(fi, fj, fmask) * C_SUB + sub];
            int4 ent = d.cFEnt[x >= 0 ? x : 0];
            while (x >= 0) {
                const int nx = ent.x;
                const int4 nent = d.cFEnt[nx >= 0 ? nx : 0];
                if (ent.z == fi && ent.w == fj) {
                    const int p = ent.y;
                    const double4 sx = d.cSegX[p];  // every listed segment is live this step
                    const double len = d.cLen[p];
                    const double2 ns = d.cNrm[p];
                    const double2 xa = make_double2(sx.x, sx.y), xb = make_double2(sx.z, sx.w);
                    const unsigned long long key = walkKey(p, xa, xb);
                    if (key != ~0ULL) consider(p, xa, xb, len, ns, key);
                }
                x = nx;
                ent = nent;
            }
        } else if (sub == 0) {
            // The reference's own walk over the nine cells.
            for (int dj = -1; dj <= 1; dj++) {
                const int j = cj + dj;
                if (j < 0 || j >= g.rows) continue;
                for (int di = -1; di <= 1; di++) {
                    const int i = ci + di;
                    if (i < 0 || i >= g.cols) continue;
                    const unsigned long long rank = (unsigned long long)((dj + 1) * 3 + (di + 1)) << 40;
                    for (int x = d.cHead[c_bucket(i, j, g.mask)]; x >= 0; x = d.cNext[x]) {
                        const int2 cell = d.cEntCell[x];
                        if (cell.x != i || cell.y != j) continue;
                        const int p = x >> 1;
                        const double4 sx = d.cSegX[p];
                        consider(p, make_double2(sx.x, sx.y), make_double2(sx.z, sx.w), d.cLen[p],
                                 d.cNrm[p], rank | (unsigned long long)x);
                    }
                }
            }
        }
        // The group's best candidate: deepest, then the earliest walk position (positions are
        // distinct, so the result does not depend on which lane found what); a fall seen by
        // any lane counts.
        int fellAny = fell ? 1 : 0;
#pragma unroll
        for (int off = 1; off < C_SUB; off <<= 1) {
            const double od = __shfl_xor(deepest, off, 64);
            const unsigned long long ok = __shfl_xor(bestKey, off, 64);
            const int ob = __shfl_xor(best, off, 64);
            fellAny |= __shfl_xor(fellAny, off, 64);
            if (od > deepest || (od == deepest && ok < bestKey)) {
                deepest = od;
                bestKey = ok;
                best = ob;
            }
        }
        fell = fellAny != 0;
        if (sub != 0) continue;
        // The previous master may have gone too deep to be captured: the drop test. It may
        // no longer be live, so its ends are read from the nodes.
        if (!c->cFOverflow && last >= 0 && !fell) {
            int a, b;
            seg_ends(d, last, a, b);
            if (walkKey(last, cpos(d, a), cpos(d, b)) != ~0ULL) {
                const double dd = c_depth(d, xn, last, a, b);
                if (dd > 0.0 && dd > C_MAX_DEPTH * d.cLen[last] && !c_excluded(d, n, pn, last, a, b)) fell = true;
            }
        }
        // accumulate :449-459
        d.cPairOk[n] = 0;
        if (best < 0) {
            if (last >= 0 && fell) escaped++;
            d.cLast[n] = -1;
            continue;
        }
        d.cLast[n] = best;

        // resolve :660-762
        int a, b;
        seg_ends(d, best, a, b);
        const double4 bx = d.cSegX[best];
        const double2 xa = make_double2(bx.x, bx.y), xb = make_double2(bx.z, bx.w);
        const double dep = c_depth_at(d, xn, best, xa, xb);
        const double len = d.cLen[best];
        const double tr0 = (xb.x - xa.x) / len, tz0 = (xb.y - xa.y) / len;
        const double tpar = ((xn.x - xa.x) * tr0 + (xn.y - xa.y) * tz0) / len;  // parameter :652-658
        const double wa = 1.0 - tpar, wb = tpar;
        const double ma = d.mass[a], mb = d.mass[b], mn = d.mass[n];
        const double invMaster = 0.25 / ma + 0.25 / mb;
        if (!(invMaster > 0.0)) continue;
        const double mMaster = 1.0 / invMaster;
        const double mStar = mn * mMaster / (mn + mMaster);
        const double k = d.cScale * mStar / (d.cRef * d.cRef);
        const double2 ns = d.cNrm[best];
        const double nR = ns.x, nZ = ns.y;
        const double mTrue = 1.0 / (wa * wa / ma + wb * wb / mb);
        const double mTangent = mn * mTrue / (mn + mTrue);
        const double2 vn = d.v[n], va = d.v[a], vb = d.v[b];
        const double relR = vn.x - wa * va.x - wb * vb.x;
        const double relZ = vn.y - wa * va.y - wb * vb.y;
        const double closing = relR * nR + relZ * nZ;
        const double damper = 2.0 * d.cDamping * mStar * __builtin_sqrt(d.cScale) / d.cRef;
        const double normal = nfd::jmax(0.0, k * dep - damper * closing);
        double* rec = d.cPair + (size_t)n * CPAIR;
        rec[PR_NORMAL] = normal;
        rec[PR_WA] = wa;
        rec[PR_WB] = wb;
        rec[PR_NR] = nR;
        rec[PR_NZ] = nZ;
        rec[PR_ENERGY] = 0.5 * k * dep * dep;
        rec[PR_DEPTH] = dep;
        const double damperForce = normal - k * dep;
        double damperLoss = 0.0;
        if (damperForce != 0.0) {
            const double closingAfter = closing + damperForce * dt / mTangent;
            damperLoss = damperForce * 0.5 * (closing + closingAfter) * dt;
        }
        rec[PR_DAMPER] = damperLoss;
        double tangent = 0.0, frictionLoss = 0.0;
        if (!(d.cFriction <= 0.0)) {
            const double tR = -nZ, tZ = nR;
            const double slide = relR * tR + relZ * tZ;
            tangent = -C_STICK_RELAXATION * mTangent * slide / dt;
            const double cone = d.cFriction * normal;
            if (tangent > cone) tangent = cone;
            else if (tangent < -cone) tangent = -cone;
            const double after = slide + tangent * dt / mTangent;
            frictionLoss = -tangent * 0.5 * (slide + after) * dt;
        }
        rec[PR_TANGENT] = tangent;
        rec[PR_FRICTION] = frictionLoss;
        d.cPairOk[n] = 1;
        d.cPairNext[n] = atomicExch(&d.cPairHead[best], n);
    }
    for (int off = 32; off > 0; off >>= 1) escaped += __shfl_xor(escaped, off, 64);
    if ((threadIdx.x & 63) == 0 && escaped) atomicAdd((unsigned long long*)&c->cEscaped, (unsigned long long)escaped);
}

// The pair terms that touch one node, added to its gathered force in the reference's
// order: by slave node ascending (accumulate walks the surface nodes in order), and within
// a pair the normal terms before the tangential ones (resolve :687-759).
// One pair term on node n: slave `who`, role 0 (n is the slave), 1 or 2 (n is end a or b of
// the master segment), resolve :687-759.
__device__ __forceinline__ void c_term(const Dev& d, double2& f, int who, int role, bool friction) {
    const double* rec = d.cPair + (size_t)who * CPAIR;
    const double normal = rec[PR_NORMAL], nR = rec[PR_NR], nZ = rec[PR_NZ];
    const double tangent = rec[PR_TANGENT];
    const double tR = -nZ, tZ = nR;
    if (role == 0) {
        f.x += normal * nR;
        f.y += normal * nZ;
        if (friction) {
            f.x += tangent * tR;
            f.y += tangent * tZ;
        }
    } else {
        const double w = role == 1 ? rec[PR_WA] : rec[PR_WB];
        f.x -= w * normal * nR;
        f.y -= w * normal * nZ;
        if (friction) {
            f.x -= w * tangent * tR;
            f.y -= w * tangent * tZ;
        }
    }
}

// Terms a node keeps in registers; a node with more is handled by c_apply_many. (A build
// with NF_CONTACT_FAST_TERMS=1 sends every node with two or more terms down that path.)
#ifndef NF_CONTACT_FAST_TERMS
#define NF_CONTACT_FAST_TERMS 8
#endif
constexpr int C_FAST_TERMS = NF_CONTACT_FAST_TERMS;

// Inserts key into t, kept ascending with unused slots at INT_MAX, by constant indices only
// (so t stays in registers).
__device__ __forceinline__ void c_insert(int (&t)[C_FAST_TERMS], int key) {
#pragma unroll
    for (int k = C_FAST_TERMS - 1; k > 0; k--) t[k] = t[k - 1] > key ? t[k - 1] : (t[k] > key ? key : t[k]);
    t[0] = t[0] > key ? key : t[0];
}

// A node with more than C_FAST_TERMS terms (rare): the terms in slave order by repeated
// selection of the next larger key over the first C_MAX_TERMS terms in collection order,
// which needs no per-thread arrays.
__device__ __forceinline__ void c_apply_many(const Dev& d, int n, bool friction, int m) {
    const int count = m < d.cMaxTerms ? m : d.cMaxTerms;
    double2 f = d.f[n];
    int prev = -1;
    for (int i = 0; i < count; i++) {
        int best = 0x7FFFFFFF;
        int c = 0;
        if (d.cPairOk[n]) {
            const int key = 4 * n;
            if (key > prev && key < best) best = key;
            c++;
        }
        for (int k = d.cNodeSegStart[n]; k < d.cNodeSegStart[n + 1] && c < count; k++) {
            const int p = d.cNodeSegOf[k];
            if (!d.cActive[p]) continue;
            int a, b;
            seg_ends(d, p, a, b);
            for (int s = d.cPairHead[p]; s >= 0 && c < count; s = d.cPairNext[s], c++) {
                const int key = 4 * s + (n == a ? 1 : 2);
                if (key > prev && key < best) best = key;
            }
        }
        c_term(d, f, best >> 2, best & 3, friction);
        prev = best;
    }
    d.f[n] = f;
}

__global__ void __launch_bounds__(BLOCK) k_c_apply(Dev d, int q, int checkGo) {
    if (c_skip(d, q, checkGo)) return;
    Ctl* c = d.ctl;
    const bool any = c->cNAct > 0;
    const bool friction = !(d.cFriction <= 0.0);  // :725, NaN included
    double acc[2] = {0.0, 0.0};
    double maxPen = 0.0;
    int pairs = 0;
    const int stride = gridDim.x * BLOCK;
    for (int n = blockIdx.x * BLOCK + threadIdx.x; any && n < d.N; n += stride) {
        if (!d.cSurfFlag[n]) continue;
        // The node's terms, keyed 4 who + role (a node meets each pair at most once, so the
        // keys sort by slave node), in the reference's order: by slave node ascending.
        int t[C_FAST_TERMS];
#pragma unroll
        for (int k = 0; k < C_FAST_TERMS; k++) t[k] = 0x7FFFFFFF;
        int m = 0;
        if (d.cPairOk[n]) {
            c_insert(t, 4 * n);
            m++;
            const double* rec = d.cPair + (size_t)n * CPAIR;
            pairs++;
            acc[0] += rec[PR_ENERGY];
            acc[1] = (acc[1] - rec[PR_DAMPER]) + rec[PR_FRICTION];
            if (rec[PR_DEPTH] > maxPen) maxPen = rec[PR_DEPTH];
        }
        for (int k = d.cNodeSegStart[n]; k < d.cNodeSegStart[n + 1]; k++) {
            const int p = d.cNodeSegOf[k];
            if (!d.cActive[p]) continue;
            int a, b;
            seg_ends(d, p, a, b);
            for (int s = d.cPairHead[p]; s >= 0; s = d.cPairNext[s]) {
                if (m == d.cMaxTerms) {  // the first cMaxTerms (C_MAX_TERMS by default) are applied
                    c->cOverflow = 1;
                    break;
                }
                if (m < C_FAST_TERMS) c_insert(t, 4 * s + (n == a ? 1 : 2));
                m++;
            }
        }
        if (m == 0) continue;
        if (m > C_FAST_TERMS) {
            c_apply_many(d, n, friction, m);
            continue;
        }
        double2 f = d.f[n];
#pragma unroll
        for (int k = 0; k < C_FAST_TERMS; k++)
            if (k < m) c_term(d, f, t[k] >> 2, t[k] & 3, friction);
        d.f[n] = f;
    }
    // Fixed-order per-block audit partials; block 0 of the nodal launch adds them up.
    __shared__ double lds[4 * 2];
    block_sum<2>(acc, lds);
    for (int off = 32; off > 0; off >>= 1) {
        pairs += __shfl_xor(pairs, off, 64);
        const double o = __shfl_xor(maxPen, off, 64);
        maxPen = o > maxPen ? o : maxPen;
    }
    __shared__ double ldsM[BLOCK / 64];
    __shared__ int ldsP[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        ldsM[threadIdx.x >> 6] = maxPen;
        ldsP[threadIdx.x >> 6] = pairs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0.0;
        int pc = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            mx = ldsM[w] > mx ? ldsM[w] : mx;
            pc += ldsP[w];
        }
        d.cPartial[3 * blockIdx.x] = acc[0];
        d.cPartial[3 * blockIdx.x + 1] = acc[1];
        d.cPartial[3 * blockIdx.x + 2] = mx;
        d.cPairsPartial[blockIdx.x] = pc;
    }
}

// The contact audits of the assembly just finished (Contact.begin resets pairs, energy and
// maxPenetration; dissipation and escapes run on), from k_c_apply's partials in block order.
__device__ void reduce_contact_partials(const Dev& d) {
    __shared__ double lds[4 * 2];
    __shared__ double ldsM[BLOCK / 64];
    __shared__ int ldsP[BLOCK / 64];
    double tot[2] = {0.0, 0.0};
    double mx = 0.0;
    int pc = 0;
    for (int b = threadIdx.x; b < d.cApplyBlocks; b += BLOCK) {
        tot[0] += d.cPartial[3 * b];
        tot[1] += d.cPartial[3 * b + 1];
        mx = d.cPartial[3 * b + 2] > mx ? d.cPartial[3 * b + 2] : mx;
        pc += d.cPairsPartial[b];
    }
    block_sum<2>(tot, lds);
    for (int off = 32; off > 0; off >>= 1) {
        pc += __shfl_xor(pc, off, 64);
        const double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        ldsM[threadIdx.x >> 6] = mx;
        ldsP[threadIdx.x >> 6] = pc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Ctl* c = d.ctl;
        double m = 0.0;
        int p = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            m = ldsM[w] > m ? ldsM[w] : m;
            p += ldsP[w];
        }
        c->cPairs = p;
        c->cEnergy = tot[0];
        c->cMaxPen = m;
        c->cDissipation += tot[1];
    }
}
