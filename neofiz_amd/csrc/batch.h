// Many independent solvers stepped together: the batched ensemble.
//
// Included by neofiz_hip.hip after the single-solver kernels and nf_solver.
//
// The reference runs an ensemble as one solver per thread (M0Cylinder.java:870-907,
// sweep/Sweep.java:77-106). Here the instances are concatenated into one set of SoA arrays
// (instance i owns nodes [nodeBase, nodeBase + N_i) and elements [elemBase, elemBase + E_i);
// its connectivity and force slots are offset accordingly) and every step is two launches
// over all live instances at once: an element launch over element tiles and a node launch
// over node tiles. A tile is a run of at most `tile` consecutive elements or nodes of one
// instance, so a workgroup reads one instance's uniforms (material, friction, wall, step
// control) as scalars, and each instance keeps its own double-buffered step control: its
// own dt, its own time, its own end test. An instance whose end test fails stops at once
// (its tiles return on the first load), and the tile tables are compacted to the live
// instances between blocks of steps, so finished instances cost nothing.
//
// Each tile runs the single solver's own per-element and per-node device functions
// (reduced_element / full_element, gather_node, nodal_node) on a copy of the solver's
// kernel arguments with that instance's uniforms patched in, so every instance performs
// exactly the arithmetic of a solo run: the batched state is bit-identical to the solo one.
//
// What a batch carries: reduced or full integration, any formulation and kinematics, one
// material per instance, rigid wall with friction, gravity, relaxation damping, mass
// scaling. Not carried (refused at creation): contact, damage, erosion, pressure edges,
// several materials in one instance. The wall's audit totals are not accumulated in a batch.
#pragma once

struct BInst {  // one instance's uniforms
    Mat mat;
    double friction, plane, cfl, fastestWave, massScale, gravityR, gravityZ, damping, qhg, thickness;
    int wall, wallFrictional, axi, finite;
    int nodeBase, nodes, elemBase, elems;
};

struct BTile {
    int inst, first, count, lead;  // lead: this tile holds the instance's first node
};

// A workgroup of BLOCK threads holds BLOCK / ts tiles of ts threads each (ts a multiple of
// the wave size, so a wave never straddles two instances); `tiles` of them are in use.
struct BBlock {
    int firstTile, ts, tiles, pad;
};

__device__ __forceinline__ Dev batch_dev(const Dev& d, const BInst& I) {
    Dev dl = d;
    dl.mat0 = I.mat;
    dl.matIndex = nullptr;
    dl.friction = I.friction;
    dl.plane = I.plane;
    dl.cfl = I.cfl;
    dl.fastestWave = I.fastestWave;
    dl.massScale = I.massScale;
    dl.gravityR = I.gravityR;
    dl.gravityZ = I.gravityZ;
    dl.damping = I.damping;
    dl.qhg = I.qhg;
    dl.thickness = I.thickness;
    dl.wall = I.wall;
    dl.wallFrictional = I.wallFrictional;
    dl.axi = I.axi;
    dl.finite = I.finite;
    return dl;
}

// This thread's tile (-1 when its slot of the workgroup is unused) and lane in it. The
// tile index is uniform across a wave, so it is moved to a scalar register and the tile,
// instance and step loads below it are scalar loads.
__device__ __forceinline__ int batch_tile(const BBlock* blocks, int& lane, int block) {
    const BBlock bb = blocks[block];
    const int slot = (int)threadIdx.x / bb.ts;
    lane = (int)threadIdx.x - slot * bb.ts;
    const int t = slot < bb.tiles ? bb.firstTile + slot : -1;
    return __builtin_amdgcn_readfirstlane(t);
}

// The element pass over element tiles (k_element's MODE_STEP / MODE_INITIAL without damage,
// erosion or pressure edges). MODE_INITIAL runs only for instances that have not stepped
// and are about to.
// One workgroup's share, element block `block` of the tile table (the tables T staged).
template <bool FULL, bool KIN>
__device__ __forceinline__ void b_element_block(const Dev& d, const BInst* insts, const BBlock* blocks,
                                                const BTile* tiles, Step* steps, int q, int mode,
                                                const nfm::Tables& T, int block) {
    int lane;
    const int ti = batch_tile(blocks, lane, block);
    if (ti < 0) return;
    const BTile t = tiles[ti];
    Step& st = steps[2 * t.inst + q];
    if (mode == MODE_STEP && !st.go) return;
    if (mode == MODE_INITIAL && (st.steps != 0 || !(st.time < st.endTime))) return;
    const Dev dl = batch_dev(d, insts[t.inst]);
    const double strainDt = mode == MODE_INITIAL ? 0.0 : st.dt;
    double edge = NFD_MAX_VALUE;
    bool sawNaN = false;
    if (lane < t.count) {
        const int e = t.first + lane;
        const int4 n = dl.conn[e];
        edge = FULL ? full_element<false, KIN>(dl, e, n, strainDt, T) : reduced_element<false, KIN>(dl, e, n, strainDt, T);
        sawNaN = edge != edge;
    }
    // Exact minimum of the shortest edge over the wave, one atomic per wave.
    unsigned long long bits = sawNaN ? MAX_BITS : (unsigned long long)nfm::bits_of(edge);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o < bits ? o : bits;
    }
    const unsigned long long anyNaN = __ballot(sawNaN);
    if ((threadIdx.x & 63) == 0) {
        if (bits != MAX_BITS) atomicMin(&st.minEdgeBits, bits);
        if (anyNaN) atomicOr(&st.nanEdge, 1);
    }
}

template <bool FULL, bool KIN>
__global__ void __launch_bounds__(BLOCK, FULL ? 2 : 3)
k_b_element(Dev d, const BInst* insts, const BBlock* blocks, const BTile* tiles, Step* steps, int q, int mode) {
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    b_element_block<FULL, KIN>(d, insts, blocks, tiles, steps, q, mode, T, blockIdx.x);
}

// The node pass over node tiles: k_gn's PHASE_START and PHASE_STEP per instance, without
// the wall's audit partials. PHASE_START gathers first exactly for the instances that have
// not stepped (their zero-interval initial assembly) and starts only instances whose end
// test passes (the host-side test of a solo run call).
template <bool WALL>
__device__ __forceinline__ void b_gn_block(const Dev& d, const BInst* insts, const BBlock* blocks, const BTile* tiles,
                                           Step* steps, int q, int phase, int block) {
    int lane;
    const int ti = batch_tile(blocks, lane, block);
    if (ti < 0) return;
    const BTile t = tiles[ti];
    const Step& in = steps[2 * t.inst + q];
    Step& out = steps[2 * t.inst + (q ^ 1)];
    const bool lead = t.lead && lane == 0;
    if (phase == PHASE_STEP && !in.go) {
        if (lead) out = in;
        return;
    }
    const Dev dl = batch_dev(d, insts[t.inst]);
    const bool runs = in.time < in.endTime && in.steps < in.stepLimit;
    const bool gather = phase == PHASE_STEP || (in.steps == 0 && runs);
    double dt = in.dt, dtPrev = in.dtPrev;
    bool next = runs;
    if (phase == PHASE_STEP && dl.finite) {
        const double minEdge = in.nanEdge ? NFD_MAX_VALUE : nfm::from_bits(in.minEdgeBits);
        dtPrev = in.dt;
        dt = dl.cfl * __builtin_sqrt(minEdge * dl.massScale) / dl.fastestWave;
    }
    if (lead) {
        Step o = in;
        o.dt = dt;
        o.dtPrev = dtPrev;
        o.go = next ? 1 : 0;
        if (next) {
            o.time = in.time + dt;
            o.steps = in.steps + 1;
        }
        o.nanEdge = 0;
        o.minEdgeBits = MAX_BITS;
        out = o;
    }
    if (lane >= t.count) return;
    const int i = t.first + lane;
    double2 fi;
    if (gather) {
        fi = gather_node(dl, i);
        dl.f[i] = fi;
    } else {
        fi = dl.f[i];
    }
    if (!next) return;
    const double dtAvg = 0.5 * (dtPrev + dt);
    const double half = 0.5 * dl.damping * dtAvg;
    const double decay = (1.0 - half) / (1.0 + half);
    const double gain = dtAvg / (1.0 + half);
    double acc[3] = {0.0, 0.0, 0.0};
    int contacts = 0, stuck = 0;
    nodal_node<WALL>(dl, i, fi, dt, dtAvg, decay, gain, acc, contacts, stuck);
}

template <bool WALL>
__global__ void __launch_bounds__(BLOCK) k_b_gn(Dev d, const BInst* insts, const BBlock* blocks, const BTile* tiles,
                                               Step* steps, int q, int phase) {
    b_gn_block<WALL>(d, insts, blocks, tiles, steps, q, phase, blockIdx.x);
}

// The persistent form: `count` steps of the whole batch in one cooperative launch. Every
// workgroup walks the element blocks, then the node blocks, with a grid-wide barrier after
// each walk (cooperative groups' grid sync), so the order of work is the two launches' order
// and every value is the same; the parity q alternates inside. The loop runs `count` times in
// every workgroup, and the launch is cooperative (the runtime refuses it unless the whole
// grid is resident), so every workgroup reaches every barrier.
template <bool FULL, bool KIN, bool WALL>
__global__ void __launch_bounds__(BLOCK, FULL ? 2 : 3)
k_b_persist(Dev d, const BInst* insts, const BBlock* eBlocks, const BTile* eTiles, int eCount,
            const BBlock* nBlocks, const BTile* nTiles, int nCount, Step* steps, int q, int count) {
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    for (int k = 0; k < count; k++) {
        for (int b = blockIdx.x; b < eCount; b += gridDim.x)
            b_element_block<FULL, KIN>(d, insts, eBlocks, eTiles, steps, q, MODE_STEP, T, b);
        grid.sync();
        for (int b = blockIdx.x; b < nCount; b += gridDim.x)
            b_gn_block<WALL>(d, insts, nBlocks, nTiles, steps, q, PHASE_STEP, b);
        grid.sync();
        q ^= 1;
    }
}

// ------------------------------------------------------------------ host side

struct nf_batch {
    hipStream_t stream = nullptr;
    std::vector<nf_solver*> solvers;
    std::vector<BInst> inst;
    std::vector<int> instTs;  // tile size per instance (threads per tile)
    Dev d{};
    BInst* instDev = nullptr;
    Step* steps = nullptr;
    BBlock* eBlocks = nullptr;
    BTile* eTiles = nullptr;
    BBlock* nBlocks = nullptr;
    BTile* nTiles = nullptr;
    int eBlockCount = 0, nBlockCount = 0;
    std::vector<uint8_t> live;  // instances not known to have stopped
    int q = 0;
    bool full = false, kin = false, wall = false;
    int N = 0, E = 0, P = 0;
    std::vector<void*> owned;
    Step* statusHost = nullptr;
    hipEvent_t statusEvent[2] = {nullptr, nullptr};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool graphDirty = true, useGraphs = true;
    bool persistent = false;  // NEOFIZ_BATCH_PERSISTENT=1: k_b_persist, one cooperative launch a block
    int persistGrid = 0;      // workgroups of that launch (all resident)
    static constexpr int GRAPH_STEPS = 32;

    ~nf_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        dropGraphs();
        for (hipEvent_t e : statusEvent)
            if (e) (void)hipEventDestroy(e);
        if (statusHost) (void)hipHostFree(statusHost);
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void dropGraphs() {
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        graphDirty = false;
    }
    template <typename T>
    int alloc(T** p, size_t count) {
        TRY(dev_alloc(p, count));
        if (*p) owned.push_back((void*)*p);
        return NF_OK;
    }

    // Tile tables over the live instances: each instance's elements (or nodes) cut into
    // tiles of its tile size, tiles of one size packed BLOCK / ts to a workgroup.
    int buildTiles(bool nodes, BBlock** blocksDev, BTile** tilesDev, int* blockCount) {
        std::vector<BTile> tiles;
        std::vector<BBlock> blocks;
        for (int ts = 64; ts <= BLOCK; ts <<= 1) {
            const int per = BLOCK / ts;
            int inBlock = 0;
            for (size_t i = 0; i < inst.size(); i++) {
                if (!live[i] || instTs[i] != ts) continue;
                const int base = nodes ? inst[i].nodeBase : inst[i].elemBase;
                const int count = nodes ? inst[i].nodes : inst[i].elems;
                for (int at = 0; at < count; at += ts) {
                    if (inBlock == 0) blocks.push_back({(int)tiles.size(), ts, 0, 0});
                    tiles.push_back({(int)i, base + at, std::min(ts, count - at), at == 0 ? 1 : 0});
                    blocks.back().tiles++;
                    inBlock = (inBlock + 1) % per;
                }
            }
        }
        *blockCount = (int)blocks.size();
        if (blocks.empty()) return NF_OK;
        HIP_TRY(hipMemcpyAsync(*blocksDev, blocks.data(), sizeof(BBlock) * blocks.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(*tilesDev, tiles.data(), sizeof(BTile) * tiles.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));  // the host vectors die here
        return NF_OK;
    }
    int rebuildTiles() {
        TRY(buildTiles(false, &eBlocks, &eTiles, &eBlockCount));
        TRY(buildTiles(true, &nBlocks, &nTiles, &nBlockCount));
        graphDirty = true;
        return NF_OK;
    }

    template <bool FULL, bool KIN>
    void launchElementT(int mode) {
        hipLaunchKernelGGL((k_b_element<FULL, KIN>), dim3(eBlockCount), dim3(BLOCK), 0, stream, d,
                           (const BInst*)instDev, (const BBlock*)eBlocks, (const BTile*)eTiles, steps, q, mode);
    }
    int launchElement(int mode) {
        if (eBlockCount == 0) return NF_OK;
        switch ((full ? 2 : 0) | (kin ? 1 : 0)) {
            case 0: launchElementT<false, false>(mode); break;
            case 1: launchElementT<false, true>(mode); break;
            case 2: launchElementT<true, false>(mode); break;
            default: launchElementT<true, true>(mode); break;
        }
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }
    int launchGN(int phase) {
        if (nBlockCount == 0) return NF_OK;
        if (wall)
            hipLaunchKernelGGL(k_b_gn<true>, dim3(nBlockCount), dim3(BLOCK), 0, stream, d, (const BInst*)instDev,
                               (const BBlock*)nBlocks, (const BTile*)nTiles, steps, q, phase);
        else
            hipLaunchKernelGGL(k_b_gn<false>, dim3(nBlockCount), dim3(BLOCK), 0, stream, d, (const BInst*)instDev,
                               (const BBlock*)nBlocks, (const BTile*)nTiles, steps, q, phase);
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }
    int enqueueStep() {
        TRY(launchElement(MODE_STEP));
        TRY(launchGN(PHASE_STEP));
        q ^= 1;
        return NF_OK;
    }
    template <bool FULL, bool KIN, bool WALL>
    int launchPersistT() {
        auto kernel = k_b_persist<FULL, KIN, WALL>;
        if (persistGrid == 0) {
            int dev = 0, perCU = 0;
            hipDeviceProp_t prop;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipGetDeviceProperties(&prop, dev));
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, BLOCK, 0));
            persistGrid = std::max(1, perCU * prop.multiProcessorCount);
        }
        const int grid = std::max(1, std::min(persistGrid, std::max(eBlockCount, nBlockCount)));
        const BInst* ip = instDev;
        const BBlock* eb = eBlocks;
        const BTile* et = eTiles;
        const BBlock* nb = nBlocks;
        const BTile* nt = nTiles;
        int ec = eBlockCount, nc = nBlockCount, qq = q, count = GRAPH_STEPS;
        void* args[] = {&d, &ip, &eb, &et, &ec, &nb, &nt, &nc, &steps, &qq, &count};
        HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kernel), dim3(grid), dim3(BLOCK), args, 0, stream));
        return NF_OK;  // an even number of steps: the parity is unchanged
    }
    int launchPersist() {
        switch ((full ? 4 : 0) | (kin ? 2 : 0) | (wall ? 1 : 0)) {
            case 0: return launchPersistT<false, false, false>();
            case 1: return launchPersistT<false, false, true>();
            case 2: return launchPersistT<false, true, false>();
            case 3: return launchPersistT<false, true, true>();
            case 4: return launchPersistT<true, false, false>();
            case 5: return launchPersistT<true, false, true>();
            case 6: return launchPersistT<true, true, false>();
            default: return launchPersistT<true, true, true>();
        }
    }
    int launchSteps() {
        if (persistent) return launchPersist();
        if (!useGraphs) {
            for (int k = 0; k < GRAPH_STEPS; k++) TRY(enqueueStep());
            return NF_OK;
        }
        if (graphDirty) dropGraphs();
        if (!graph[q]) {
            hipGraph_t g = nullptr;
            const int q0 = q;
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            int rc = NF_OK;
            for (int k = 0; k < GRAPH_STEPS && rc == NF_OK; k++) rc = enqueueStep();
            hipError_t err = hipStreamEndCapture(stream, &g);
            q = q0;
            if (rc) return rc;
            if (err != hipSuccess) return fail(NF_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(err));
            err = hipGraphInstantiate(&graph[q], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) return fail(NF_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(err));
        }
        HIP_TRY(hipGraphLaunch(graph[q], stream));
        return NF_OK;  // an even number of steps: the parity is unchanged
    }

    // Every instance runs while its time is below end_time (and, when max_steps > 0, for at
    // most max_steps steps in this call), as a solo runUntil would.
    int runUntil(double endTime, long long maxSteps) {
        const int I = (int)inst.size();
        std::vector<Step> h(2 * I);
        HIP_TRY(hipMemcpyAsync(h.data(), steps, sizeof(Step) * 2 * I, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for (int i = 0; i < I; i++) {
            Step& s = h[2 * i + q];
            s.endTime = endTime;
            s.stepLimit = maxSteps > 0 ? s.steps + maxSteps : 0x7FFFFFFFFFFFFFFFLL;
            live[i] = s.time < endTime && maxSteps != 0;
        }
        HIP_TRY(hipMemcpyAsync(steps, h.data(), sizeof(Step) * 2 * I, hipMemcpyHostToDevice, stream));
        TRY(rebuildTiles());  // synchronises: h may go
        if (eBlockCount == 0) return NF_OK;  // nothing to run
        TRY(launchElement(MODE_INITIAL));
        TRY(launchGN(PHASE_START));
        q ^= 1;
        if (!statusHost) {
            HIP_TRY(hipHostMalloc((void**)&statusHost, 2 * sizeof(Step) * I, hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&statusEvent[0], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&statusEvent[1], hipEventDisableTiming));
        }
        // Blocks of steps with one block of look-ahead; at each status check, instances
        // seen stopped leave the tile tables (a stopped instance never restarts).
        int slot = 0;
        bool pending = false;
        while (true) {
            TRY(launchSteps());
            HIP_TRY(hipMemcpy2DAsync(statusHost + slot * I, sizeof(Step), steps + q, 2 * sizeof(Step), sizeof(Step), I,
                                     hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(statusEvent[slot], stream));
            if (pending) {
                HIP_TRY(hipEventSynchronize(statusEvent[slot ^ 1]));
                const Step* st = statusHost + (slot ^ 1) * I;
                int alive = 0, dropped = 0;
                for (int i = 0; i < I; i++) {
                    if (!live[i]) continue;
                    if (st[i].go) alive++;
                    else {
                        live[i] = 0;
                        dropped++;
                    }
                }
                if (alive == 0) break;
                if (dropped) {
                    HIP_TRY(hipStreamSynchronize(stream));
                    TRY(rebuildTiles());
                }
            }
            pending = true;
            slot ^= 1;
        }
        HIP_TRY(hipStreamSynchronize(stream));
        return NF_OK;
    }
};

// Copies `count` elements of T between a solver's buffer and the batch's, at an offset.
template <typename T>
static int batch_copy(T* dst, const T* src, size_t count, hipStream_t stream) {
    if (count == 0) return NF_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToDevice, stream));
    return NF_OK;
}

// Moves every per-node and per-point array between the batch and its solvers.
static int batch_transfer(nf_batch* b, bool toBatch) {
    for (size_t i = 0; i < b->solvers.size(); i++) {
        nf_solver* s = b->solvers[i];
        // the solver's control block and state change underneath it (back from the batch)
        s->touch();
        s->fastOpen = false;
        const BInst& I = b->inst[i];
        const size_t nb = I.nodeBase, eb = I.elemBase, pb = (size_t)I.elemBase * s->ppe;
        const size_t N = s->N, E = s->E, P = s->P;
        Dev& d = b->d;
        Dev& o = s->d;
#define MOVE(field, base, n) TRY(toBatch ? batch_copy(d.field + base, o.field, n, b->stream) : batch_copy(o.field, d.field + base, n, b->stream))
        MOVE(u, nb, N);
        MOVE(v, nb, N);
        MOVE(f, nb, N);
        MOVE(flags, nb, N);
        MOVE(sig0, pb, P);
        MOVE(sig1, pb, P);
        MOVE(sig2, pb, P);
        MOVE(sig3, pb, P);
        MOVE(back0, pb, P);
        MOVE(back1, pb, P);
        MOVE(back2, pb, P);
        MOVE(back3, pb, P);
        MOVE(epsP, pb, P);
        MOVE(work, pb, P);
        MOVE(hgR, eb, E);
        MOVE(hgZ, eb, E);
        MOVE(hgK, eb, E);
#undef MOVE
        if (toBatch) {
            TRY(batch_copy((double2*)d.x0 + nb, o.x0, N, b->stream));
            TRY(batch_copy((double*)d.z0 + nb, o.z0, N, b->stream));
            TRY(batch_copy((double*)d.mass + nb, o.mass, N, b->stream));
        }
        // The step control: the solver's current buffer into both of the batch's, or the
        // batch's current buffer back into both of the solver's.
        for (int k = 0; k < 2; k++) {
            if (toBatch) HIP_TRY(hipMemcpyAsync(b->steps + 2 * i + k, &s->ctl->s[s->q], sizeof(Step), hipMemcpyDeviceToDevice, b->stream));
            else HIP_TRY(hipMemcpyAsync(&s->ctl->s[k], b->steps + 2 * i + b->q, sizeof(Step), hipMemcpyDeviceToDevice, b->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return NF_OK;
}
