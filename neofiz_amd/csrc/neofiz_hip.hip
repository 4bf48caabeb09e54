// libneofiz_hip -- Neofiz's explicit step loop (solver/ExplicitSolver.java:940-1115) as
// hand-written gfx950 FP64 kernels behind the C ABI of include/neofiz.h.
//
// One step is three launches on the handle's stream:
//   k_nodal    central-difference update of every node, rigid anvil with Coulomb cone
//              (solver/ExplicitSolver.java:998-1062, solver/RigidWall.java:163-279);
//   k_element  erosion test, then the element pass: mid-step kinematics, Cayley co-rotation,
//              J2 / Johnson-Cook return map, damage onset, internal and hourglass forces,
//              written to per-element force slots (:1219-1504); shortest edge by an exact
//              atomic minimum (min is associative, so the result is order-free);
//   k_gather   fixed-order CSR sum of each node's slots plus gravity (:1147-1165), and the
//              adaptive CFL step (:985-990).
// Every control scalar (time, dt, step count, the while-time-below-end guard) lives on the
// device, so steps chain without a host round trip and a block of steps can be captured in a
// hipGraph. No atomic touches a force or a state value: forces are summed in the reference's
// order, so the GPU reproduces the reference arithmetic bit for bit. Audit accumulators that
// the reference sums serially (wall impulse, eroded energy) are reduced per step in a fixed
// order instead; they agree to rounding, not bitwise.
//
// Layout (structure of arrays, 16-byte pairs where a thread reads both components):
//   nodes     x0[N] (r, z), u[N] (ur, uz), v[N] (vr, vz), f[N] (fr, fz), mass[N], flags[N]
//   elements  conn[E] int4, matIndex[E], gone[E]
//   points    sig0..3[P], back0..3[P], epsP[P], work[P]; hgR, hgZ, hgK [E]
//   slots     slot[4 E + 2 edges] (gr, gz); CSR slotStart[N + 1], slotOf[4 E + 2 edges]
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/neofiz.h"
#include "device_physics.h"

using nfd::Mat;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t err_ = (expr);                                                      \
        if (err_ != hipSuccess) {                                                      \
            return fail(err_ == hipErrorOutOfMemory ? NF_E_MEMORY : NF_E_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(err_));          \
        }                                                                              \
    } while (0)

constexpr int BLOCK = 256;
// Nodes of one patch the fused step kernel holds in LDS (a lattice patch of 8 x 8 elements has 81).
constexpr int STEP_PATCH_NODES = 128;
constexpr uint8_t FIXED_R = 1, FIXED_Z = 2, ON_WALL = 4;
constexpr int MODE_STEP = 0, MODE_INITIAL = 1, MODE_ADVANCE = 2;

// Device-resident control. The per-step scalars are double-buffered: a launch reads buffer
// q and block 0 writes buffer q ^ 1, so no block ever reads a field that another block of
// the same launch writes. The accumulators below them have one writer per launch.
struct Step {
    double time, dt, dtPrev, endTime;
    long long steps, stepLimit;
    int go;       // the element pass that reads this buffer executes
    int nanEdge;  // some element's shortest edge was NaN
    unsigned long long minEdgeBits;  // exact minimum of the squared shortest edge, as bits
    // The fused step kernel (k_step): which of the two node-state and slot buffers holds the
    // current state (0: the canonical arrays), and whether the launch that wrote this buffer
    // left wall audit partials (in wallPartialF[this buffer's index]) for the next to add.
    int buf;
    int pendingWall;
};

struct Ctl {
    Step s[2];
    int wallContact, wallStuck;
    double wallImpulse, wallEnergyLoss, wallFriction;
    int erodedCount, invertedCount;
    double erodedEnergy;
    int surfaceStale;
    int pad;
    // Penalty contact (solver/Contact.java). Counts of the compact lists of live surface
    // segments and surface nodes; the grid's bounding box as order-preserving keys
    // {rlo, zlo, -rhi, -zhi} (all minima) and the longest segment's bits; the audits.
    int cNAct, cNSurf;
    unsigned long long cBox[4];
    unsigned long long cLongest;
    int cPairs, cOverflow;
    double cEnergy, cDissipation, cMaxPen;
    long long cEscaped;
    int cFCount, cFOverflow;  // entries of the capture grid this assembly; it overflowed
    // The fused step kernel's shortest-edge minima, rotated over three slots: launch k reads
    // slot k % 3, atomically lowers slot (k + 1) % 3 and resets slot (k + 2) % 3, so that no
    // slot is reset while another workgroup of the same launch reads or lowers it.
    unsigned long long fMin[3];
    int fNaN[3];
    // The step count at the start of the current run call (k_open_run), for its step tally.
    long long runStart;
};

constexpr unsigned long long MAX_BITS = 0x7FEFFFFFFFFFFFFFULL;  // Double.MAX_VALUE

// Everything a kernel needs, passed by value.
struct Dev {
    Ctl* ctl;
    int N, E, P, edges, ppe;
    const double2* x0;
    const double* z0;  // axial reference coordinate alone, for the wall test in the node pass
    double2* u;
    double2* v;
    double2* f;
    const double* mass;
    uint8_t* flags;
    const int4* conn;
    // The element pass's schedule (null order: thread t takes element t). Elements are grouped
    // into patches of 64 neighbours, one wave each; wave slot s takes patch patchOf[s] (-1: no
    // patch), lane l element order[64 p + l] (-1 past the last element).
    const int* order;     // [64 patches]
    const int4* orderConn;  // [64 patches] the corners of order[k]
    // Both double-buffered by the element pass's control-buffer index q: the launch that reads
    // Step q takes its schedule from patchOf + q * 4 patchBlocks and writes its costs to
    // patchCost + q * patches.
    int* patchOf;         // [2][4 patchBlocks]
    uint8_t* patchCost;   // [2][patches] Newton cost of the patch's slowest lane, last step
    int patches, patchBlocks, units;  // units: the device's compute units
    // The fused step kernel (k_step). Patch p's nodes are entries pnStart[p] .. pnStart[p+1]
    // of pnCode (ascending node numbers): pnCode holds the node number in bits 0-27, the
    // node's force-slot count less one in bits 28-29, and bit 31 set when p owns the node (p
    // holds the node's lowest-numbered element). orderLocal[64 p + l] are the corners of element
    // order[64 p + l] as indices into the patch's list. Node state and force slots are
    // double-buffered: buffer 0 is u, v, flags, slotN; buffer 1 is uB, vB, flagsB, slotNB.
    // The fused run's force slots are node-major: slotN[4 n + r] is node n's r-th slot in
    // the gather's order (every node has at most four), so that a node reads its slots as one
    // 64-byte record; orderDest[64 p + l] gives element order[64 p + l]'s four corner slots'
    // places there. The two-launch step keeps the element-major d.slot.
    const int* pnStart;
    // patchRun[s] = {patchOf[s], pnStart[p], pnStart[p+1] - pnStart[p], 0} for p = patchOf[s]
    // (or {-1, 0, 0, 0}), double-buffered as patchOf: one scalar load gives a fused-step wave
    // its patch and node list.
    int4* patchRun;
    const int* pnCode;
    const uchar4* orderLocal;
    const int4* orderDest;
    double2* uB;
    double2* vB;
    uint8_t* flagsB;
    double2* slotN;
    double2* slotNB;
    double* wallPartialF;     // [2][patchBlocks * 3]
    int* wallCountPartialF;   // [2][patchBlocks * 2]
    const int* matIndex;  // null when uniform
    const Mat* mats;
    Mat mat0;             // the uniform material, by value (kernel arguments sit in SGPRs)
    uint8_t* gone;        // null when erosion is off
    double* sig0; double* sig1; double* sig2; double* sig3;
    double* back0; double* back1; double* back2; double* back3;
    double* epsP; double* work;
    double* hgR; double* hgZ; double* hgK;
    double2* slot;
    const int* slotStart;
    const int4* slot4;  // per node: its first four slots in gather order (-1 unused)
    const int* slotOf;
    const int2* pedges;
    // damage (null when off)
    double* dStrength; double* dSoftening; double* dWidth;
    const double* dFailure; const double* dGf; const double* dYoungs; const double* dGeom;
    double erosionDamage;
    // per-block partial sums for the ordered reductions
    double* wallPartial;
    int* wallCountPartial;
    double* erodePartial;
    // uniforms
    int axi, finite, wall, wallFrictional;
    double thickness, qhg, damping, cfl, fastestWave, massScale;
    double gravityR, gravityZ;
    double plane, friction;  // friction may be +inf
    double targetPressure, rampDuration;
    int nodalBlocks, elementBlocks;
    // Penalty contact (null when off). Segments are indexed by their potential index
    // p = 4 e + k (edge k of element e, from corner k to corner k + 1), which orders them
    // exactly as the reference's surface list does; surface nodes by node index.
    int contact;
    uint8_t* cDead;          // [E] elements the contact has forgotten
    uint8_t* cActive;        // [4E] p is a live surface segment
    uint8_t* cSurfFlag;      // [N] node is on the surface
    uint8_t* cPairOk;        // [N] node resolved a pair this assembly
    int* cAct;               // compact list of live segments (any order)
    int* cSurf;              // compact list of surface nodes (any order)
    const int* cShareStart;  // [4E+1] CSR: the other potential segments with the same edge
    const int* cShareOf;
    const int* cNodeSegStart;  // [N+1] CSR: potential segments touching a node, ascending
    const int* cNodeSegOf;
    const int* cPartner;       // [4E] the one other potential segment with the same edge, -1
                               // if none, -2 if several (then the CSR above is walked)
    const int4* cNodeSeg8;     // [2N] a node's first eight CSR entries, -1 padded
    double* cLen;            // [4E]
    double2* cNrm;           // [4E] outward unit normal
    double2* cPn;            // [N] node normal
    int* cLast;              // [N] lastMaster (potential index) or -1
    int* cHead;              // hash-grid heads
    int* cNext;              // [8E] per entry 2p + end
    int2* cEntCell;          // [8E] the entry's grid cell
    int* cPairHead;          // [4E] resolved slaves whose master is p
    int* cPairNext;          // [N]
    double* cPair;           // [N * CPAIR] the resolved pair of each slave node
    double* cPartial;        // [cApplyBlocks * 3] energy, dissipation, max penetration
    int* cPairsPartial;      // [cApplyBlocks]
    int cApplyBlocks;
    double cFriction, cDamping, cScale, cRef;
    // The capture grid: square cells of side cFine, each listing every segment whose capture
    // rectangle (the region where it can be a node's master) overlaps the cell.
    int* cFHead;             // [cFHashMax * C_SUB] hash heads, C_SUB lists per bucket
    int4* cFEnt;             // [cFCap] {next entry, segment p, cell i, cell j}: one load a link
    double4* cSegX;          // [4E] current end positions of each live segment (a.r, a.z, b.r, b.z)
    int cFCap;
    unsigned cFHashMax;
    double cFine;
    int cMaxTerms;           // pair terms one node may take in one step (C_MAX_TERMS unless lowered)
};

__device__ __forceinline__ double pressureAt(const Dev& d, double t) {  // :1203-1207
    if (d.rampDuration <= 0.0) return d.targetPressure;
    double s = nfd::jmin(1.0, t / d.rampDuration);
    return d.targetPressure * s * s * (3.0 - 2.0 * s);
}

// Fixed-order block sum of up to 8 doubles (wave shuffle tree, then LDS in wave order).
template <int K>
__device__ void block_sum(double (&vals)[K], double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = vals[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        vals[k] = x;
    }
    if (lane == 0)
        for (int k = 0; k < K; k++) lds[wave * K + k] = vals[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int w = 0; w < BLOCK / 64; w++) s += lds[w * K + k];
            vals[k] = s;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ nodal update
// One node of the central-difference update, solver/ExplicitSolver.java:998-1062, with the
// rigid wall of solver/RigidWall.java:163-279. acc receives this node's contribution to the
// wall's impulse, arrival energy loss and friction dissipation.
// The central-difference update of one node (solver/ExplicitSolver.java:998-1062, with the
// rigid wall of solver/RigidWall.java:163-279) on values in registers: mass m, reference
// axial coordinate z0, flags fl, velocity vi, displacement ui, gathered force fi. The caller
// loads and stores them (nodal_node below, or the fused step kernel's node phase).
template <bool WALL>
__device__ __forceinline__ void nodal_core(const Dev& d, double m, double z0, uint8_t& fl, double2& vi,
                                           double2& ui, double2 fi, double dt, double dtAvg,
                                           double decay, double gain, double (&acc)[3], int& contacts,
                                           int& stuck) {
    const double invM = 1.0 / m;
    const bool resting = WALL && (fl & ON_WALL) && fi.y <= 0.0;
    if (fl & FIXED_R) {
        vi.x = 0.0;
    } else if (WALL && resting && d.wallFrictional) {
        // RigidWall.slide :252-279
        const double normal = -fi.y;
        const double trial = -m * decay * vi.x / gain - fi.x;
        const double limit = __builtin_isinf(d.friction) ? HUGE_VAL : d.friction * normal;
        if (nfd::jabs(trial) <= limit) {
            acc[1] += 0.5 * m * vi.x * vi.x;
            stuck++;
            vi.x = 0.0;
        } else {
            const double tangential = __builtin_copysign(limit, trial);
            const double slid = decay * vi.x + gain * (fi.x + tangential) / m;
            acc[2] += -tangential * slid * dt;
            vi.x = slid;
        }
        ui.x += vi.x * dt;
    } else {
        vi.x = decay * vi.x + gain * fi.x * invM;
        ui.x += vi.x * dt;
    }
    if (fl & FIXED_Z) {
        vi.y = 0.0;
    } else if (WALL && resting) {
        acc[0] += -fi.y * dtAvg;  // RigidWall.hold :175-178
        contacts++;
    } else {
        const double trial = decay * vi.y + gain * fi.y * invM;
        const double zc = z0 + ui.y;
        if (WALL && zc + trial * dt < d.plane) {
            acc[0] += -m * trial;  // RigidWall.arrive :201-205
            acc[1] += 0.5 * m * trial * trial;
            contacts++;
            ui.y += d.plane - zc;
            vi.y = 0.0;
            fl |= ON_WALL;
        } else {
            vi.y = trial;
            ui.y += vi.y * dt;
            if (WALL) fl &= (uint8_t)~ON_WALL;
        }
    }
}

template <bool WALL>
__device__ __forceinline__ void nodal_node(const Dev& d, int i, double2 fi, double dt,
                                           double dtAvg, double decay, double gain,
                                           double (&acc)[3], int& contacts, int& stuck) {
    uint8_t fl = d.flags[i];
    double2 vi = d.v[i];
    double2 ui = d.u[i];
    nodal_core<WALL>(d, d.mass[i], WALL ? d.z0[i] : 0.0, fl, vi, ui, fi, dt, dtAvg, decay, gain, acc, contacts, stuck);
    d.v[i] = vi;
    d.u[i] = ui;
    if (WALL) d.flags[i] = fl;
}

// Gather of one node, solver/ExplicitSolver.java:1147-1165: the node's force slots in the
// reference's order (elements ascending, then pressure edges), plus gravity.
// A node's first four slots come from one 16-byte load, so their four loads issue together
// instead of each waiting on an index load; a node with more sources (rare on a quad mesh)
// continues through the CSR. The sum runs in the same order either way.
// slot: the force slots to gather (d.slot, or the fused step's other slot buffer).
__device__ __forceinline__ double2 gather_slots(const Dev& d, const double2* __restrict__ slot, int n) {
    const int4 k4 = d.slot4[n];
    const double2 z = make_double2(0.0, 0.0);
    const double2 s0 = k4.x >= 0 ? slot[k4.x] : z;
    const double2 s1 = k4.y >= 0 ? slot[k4.y] : z;
    const double2 s2 = k4.z >= 0 ? slot[k4.z] : z;
    const double2 s3 = k4.w >= 0 ? slot[k4.w] : z;
    double sr = 0.0, sz = 0.0;
    if (k4.x >= 0) { sr += s0.x; sz += s0.y; }
    if (k4.y >= 0) { sr += s1.x; sz += s1.y; }
    if (k4.z >= 0) { sr += s2.x; sz += s2.y; }
    if (k4.w >= 0) { sr += s3.x; sz += s3.y; }
    const int start = d.slotStart[n] + 4, end = d.slotStart[n + 1];
    for (int k = start; k < end; k++) {
        const double2 s = slot[d.slotOf[k]];
        sr += s.x;
        sz += s.y;
    }
    const double m = d.mass[n];
    return make_double2(sr + m * d.gravityR, sz + m * d.gravityZ);
}
__device__ __forceinline__ double2 gather_node(const Dev& d, int n) { return gather_slots(d, d.slot, n); }
// The same gather from a fused run's node-major slots (slotN, or slotNB for buffer 1: at most
// four per node, in order).
__device__ __forceinline__ double2 gather_node_major(const Dev& d, int n, int buf) {
    const int slots = d.slotStart[n + 1] - d.slotStart[n];
    const double2* a = (buf ? d.slotNB : d.slotN) + 4 * (size_t)n;
    const double2 s0 = a[0], s1 = a[1], s2 = a[2], s3 = a[3];
    double sr = 0.0, sz = 0.0;
    sr += s0.x; sz += s0.y;
    if (slots > 1) { sr += s1.x; sz += s1.y; }
    if (slots > 2) { sr += s2.x; sz += s2.y; }
    if (slots > 3) { sr += s3.x; sz += s3.y; }
    const double m = d.mass[n];
    return make_double2(sr + m * d.gravityR, sz + m * d.gravityZ);
}

// The element pass's schedule from the patch costs (defined below; the contact search's extra
// workgroup makes it when contact is on).
__device__ void schedule_patches(const Dev& d, const uint8_t* __restrict__ costIn, int* __restrict__ patchOfOut,
                                 int4* __restrict__ runOut);
#include "device_contact.h"

constexpr int PHASE_STEP = 0, PHASE_START = 1, PHASE_ASSEMBLE_END = 2, PHASE_RESUME = 3;
// With penalty contact the launch is split around the contact launches: PART_GATHER only
// gathers, PART_REST does everything after the gather.
constexpr int PART_ALL = 0, PART_GATHER = 1, PART_REST = 2;

// The node-centred launch that closes step n and opens step n + 1 (the reference's
// advanceMaterialAndAssemble tail, the adaptive step, and the next step's nodal update,
// solver/ExplicitSolver.java:947-990, :1092, :1111-1114):
//   PHASE_STEP   gather step n's forces, take the CFL step from step n's shortest edge, test
//                while-time-below-end, and if the next step runs, integrate every node;
//   PHASE_START  the first nodal update of a run call (gathering first after the initial
//                zero-interval assembly when gatherFirst is set);
//   PHASE_ASSEMBLE_END  gather only, after advance();
//   PHASE_RESUME  PHASE_START without a gather, if the step's end test passes (the adjoint's
//                forward enqueues its chunks ahead of knowing whether the run has ended).
// Wave slot of the patch ranked r (0 = costliest) in a grid of patchBlocks workgroups of
// four waves: ranks fill workgroups four at a time, and the workgroups are dealt to the compute
// units in rounds that alternate direction, so that a unit holding one of the costliest
// workgroups of the first round holds one of the cheapest of the second.
__host__ __device__ __forceinline__ int patch_slot(int r, int patchBlocks, int units) {
    const int b = r >> 2, round = b / units, k = b - round * units;
    const int inRound = patchBlocks - round * units < units ? patchBlocks - round * units : units;
    const int pos = round * units + ((round & 1) ? inRound - 1 - k : k);
    return 4 * pos + (r & 3);
}

// The element pass's schedule for a later launch, from each patch's cost in a finished one
// (costs in, schedule out): a counting sort of the patches by cost, costliest first, into
// patch_slot order. One workgroup does it. Patches of equal cost are ranked in the order the
// LDS atomics give: the schedule decides only which wave takes a patch, never a result.
__device__ void schedule_patches(const Dev& d, const uint8_t* __restrict__ costIn, int* __restrict__ patchOfOut,
                                 int4* __restrict__ runOut) {
    constexpr int CHUNK = 4096;  // costs staged in LDS per pass, read 16 at a time
    __shared__ int start[256];
    __shared__ uint4 staged[CHUNK / 16];
    const uint8_t* cost = (const uint8_t*)staged;
    const int c = threadIdx.x;  // BLOCK = 256 cost bins
    const int P = d.patches;
    const bool whole = P <= CHUNK;  // then one staging serves both passes
    auto stage = [&](int from) {
        const int n = min(CHUNK, P - from);
        __syncthreads();
        for (int k = threadIdx.x; k < (n + 15) / 16; k += BLOCK) {
            if (16 * k + 16 <= n && ((from & 15) == 0)) {
                staged[k] = *(const uint4*)(costIn + from + 16 * k);
            } else {
                for (int j = 0; j < 16; j++)
                    ((uint8_t*)staged)[16 * k + j] = 16 * k + j < n ? costIn[from + 16 * k + j] : 0;
            }
        }
        __syncthreads();
        return n;
    };
    start[c] = 0;
    for (int from = 0; from < P; from += CHUNK) {
        const int n = stage(from);
        for (int p = threadIdx.x; p < n; p += BLOCK) atomicAdd(&start[cost[p]], 1);
    }
    __syncthreads();
    // suffix sums over the bins (Hillis-Steele), then the exclusive start of each bin
    const int count = start[c];
    for (int off = 1; off < 256; off <<= 1) {
        const int add = c + off < 256 ? start[c + off] : 0;
        __syncthreads();
        start[c] += add;
        __syncthreads();
    }
    const int first = start[c] - count;
    __syncthreads();
    start[c] = first;
    for (int from = 0; from < P; from += CHUNK) {
        const int n = whole ? (__syncthreads(), P) : stage(from);
        for (int p = threadIdx.x; p < n; p += BLOCK) {
            const int r = atomicAdd(&start[cost[p]], 1);
            const int slot = patch_slot(r, d.patchBlocks, d.units);
            patchOfOut[slot] = from + p;
            if (runOut) {
                const int b0 = d.pnStart[from + p], b1 = d.pnStart[from + p + 1];
                runOut[slot] = make_int4(from + p, b0, b1 - b0, 0);
            }
        }
    }
}

__device__ void reduce_wall_partials(const Dev& d, const double* __restrict__ partial,
                                     const int* __restrict__ countPartial, int blocks);

// Reads control buffer q; block 0 writes buffer q ^ 1.
// minSlot >= 0: the preceding element pass was the fused step kernel's, whose shortest edge is
// in Ctl::fMin[minSlot].
// nodeMajor: the gather closes a fused run, whose slots are node-major, in the buffer the Step
// names (slotN, slotNB). It also brings the node state back to buffer 0 (u, v, flags) and adds
// the last fused launch's wall partials to the totals, which leaves the closed state every
// other kernel and the host read.
template <bool WALL>
__global__ void __launch_bounds__(BLOCK) k_gn(Dev d, int q, int phase, int gatherFirst, int part, int minSlot,
                                              int nodeMajor, int schedule) {
    if (schedule && blockIdx.x == gridDim.x - 1) {
        // the workgroup past the others (launchGNT): the schedule of the element launch that
        // reads Step q ^ 1, from the costs of the one that read Step q
        schedule_patches(d, d.patchCost + (size_t)q * d.patches, d.patchOf + (size_t)(q ^ 1) * 4 * d.patchBlocks,
                         d.patchRun ? d.patchRun + (size_t)(q ^ 1) * 4 * d.patchBlocks : nullptr);
        return;
    }
    Ctl* c = d.ctl;
    const Step& in = c->s[q];
    Step& out = c->s[q ^ 1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int i0 = blockIdx.x * BLOCK + threadIdx.x;
    if (nodeMajor && in.buf && i0 < d.N) {  // buffer 1 back to buffer 0, each thread its node
        d.u[i0] = d.uB[i0];
        d.v[i0] = d.vB[i0];
        d.flags[i0] = d.flagsB[i0];
    }
    if (phase == PHASE_STEP && !in.go) {  // nothing ran: carry the state over
        if (lead && part != PART_GATHER) {
            Step o = in;
            o.buf = 0;
            o.pendingWall = 0;
            out = o;
        }
        return;
    }
    const bool gathered = phase == PHASE_STEP || gatherFirst;
    const bool gather = gathered && part != PART_REST;
    if (part == PART_GATHER) {
        if ((int)blockIdx.x >= d.nodalBlocks) {  // the workgroups past the nodes: the contact rebuild
            c_rebuild(d, blockIdx.x - d.nodalBlocks);
            return;
        }
        const int i = blockIdx.x * BLOCK + threadIdx.x;
        if (gather && i < d.N) d.f[i] = nodeMajor ? gather_node_major(d, i, in.buf) : gather_node(d, i);
        return;
    }
    if (nodeMajor && WALL && in.pendingWall && blockIdx.x == 0)
        reduce_wall_partials(d, d.wallPartialF + (size_t)q * 3 * d.patchBlocks,
                             d.wallCountPartialF + (size_t)q * 2 * d.patchBlocks, d.patchBlocks);
    if (part == PART_REST && gathered && blockIdx.x == 0) reduce_contact_partials(d);
    if (phase == PHASE_STEP && d.gone && blockIdx.x == 0) {
        // eroded energy of step n: the element launch's per-patch (or, without a schedule,
        // per-block) partials, summed by this workgroup in a fixed order (strided per thread,
        // then the block tree)
        __shared__ double ldsE[4 * 1];
        double tot[1] = {0.0};
        const int parts = d.order ? d.patches : d.elementBlocks;
        for (int b = threadIdx.x; b < parts; b += BLOCK) tot[0] += d.erodePartial[b];
        block_sum<1>(tot, ldsE);
        if (threadIdx.x == 0) c->erodedEnergy += tot[0];
    }
    double dt = in.dt, dtPrev = in.dtPrev;
    bool next = true;
    if (phase == PHASE_STEP) {
        if (d.finite) {
            const bool nanEdge = minSlot >= 0 ? c->fNaN[minSlot] != 0 : in.nanEdge != 0;
            const unsigned long long bits = minSlot >= 0 ? c->fMin[minSlot] : in.minEdgeBits;
            const double minEdge = nanEdge ? NFD_MAX_VALUE : nfm::from_bits(bits);
            dtPrev = in.dt;
            dt = d.cfl * __builtin_sqrt(minEdge * d.massScale) / d.fastestWave;
        }
        next = in.time < in.endTime && in.steps < in.stepLimit;
    } else if (phase == PHASE_ASSEMBLE_END) {
        next = false;
    } else if (phase == PHASE_RESUME) {
        // PHASE_START whose end test the device makes (a chunk the host enqueued ahead)
        next = in.time < in.endTime && in.steps < in.stepLimit;
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
        o.buf = 0;
        o.pendingWall = 0;
        out = o;
    }

    const int i = blockIdx.x * BLOCK + threadIdx.x;
    double2 fi = make_double2(0.0, 0.0);
    if (i < d.N) {
        if (gather) {
            fi = nodeMajor ? gather_node_major(d, i, in.buf) : gather_node(d, i);
            d.f[i] = fi;
        } else {
            fi = d.f[i];
        }
    }
    if (!next) return;
    const double dtAvg = 0.5 * (dtPrev + dt);
    const double half = 0.5 * d.damping * dtAvg;
    const double decay = (1.0 - half) / (1.0 + half);
    const double gain = dtAvg / (1.0 + half);
    double acc[3] = {0.0, 0.0, 0.0};
    int contacts = 0, stuck = 0;
    if (i < d.N) nodal_node<WALL>(d, i, fi, dt, dtAvg, decay, gain, acc, contacts, stuck);
    if (!WALL) return;

    // Per-block wall audit partials (fixed-order block sum); the next element launch adds
    // them up in block order.
    __shared__ double lds[4 * 3];
    __shared__ int ldsi[2 * 4];
    block_sum<3>(acc, lds);
    int a = contacts, b = stuck;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        ldsi[(threadIdx.x >> 6) * 2] = a;
        ldsi[(threadIdx.x >> 6) * 2 + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* p = d.wallPartial + 3 * blockIdx.x;
        p[0] = acc[0];
        p[1] = acc[1];
        p[2] = acc[2];
        d.wallCountPartial[2 * blockIdx.x] = ldsi[0] + ldsi[2] + ldsi[4] + ldsi[6];
        d.wallCountPartial[2 * blockIdx.x + 1] = ldsi[1] + ldsi[3] + ldsi[5] + ldsi[7];
    }
}

// Adds the wall partials of the preceding nodal update to the running totals, in block
// order through a fixed tree (run by block 0 of the element launch that follows it).
__device__ void reduce_wall_partials(const Dev& d, const double* __restrict__ partial,
                                     const int* __restrict__ countPartial, int blocks) {
    __shared__ double lds[4 * 3];
    __shared__ int ldsi[2 * 4];
    double tot[3] = {0.0, 0.0, 0.0};
    int cc = 0, cs = 0;
    for (int b = threadIdx.x; b < blocks; b += BLOCK) {
        tot[0] += partial[3 * b];
        tot[1] += partial[3 * b + 1];
        tot[2] += partial[3 * b + 2];
        cc += countPartial[2 * b];
        cs += countPartial[2 * b + 1];
    }
    block_sum<3>(tot, lds);
    for (int off = 32; off > 0; off >>= 1) {
        cc += __shfl_xor(cc, off, 64);
        cs += __shfl_xor(cs, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        ldsi[(threadIdx.x >> 6) * 2] = cc;
        ldsi[(threadIdx.x >> 6) * 2 + 1] = cs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Ctl* c = d.ctl;
        c->wallImpulse += tot[0];
        c->wallEnergyLoss += tot[1];
        c->wallFriction += tot[2];
        c->wallContact = ldsi[0] + ldsi[2] + ldsi[4] + ldsi[6];
        c->wallStuck = ldsi[1] + ldsi[3] + ldsi[5] + ldsi[7];
    }
}

// ------------------------------------------------------------------ element pass

__device__ __forceinline__ Mat load_mat(const Dev& d, int e) {
    if (d.matIndex) return d.mats[d.matIndex[e]];
    return d.mat0;
}

// Signed current area, solver/ExplicitSolver.java:681-691.
__device__ __forceinline__ double currentArea(const Dev& d, int4 n) {
    const int idx[4] = {n.x, n.y, n.z, n.w};
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = idx[k], j = idx[(k + 1) & 3];
        const double2 xi = d.x0[i], ui = d.u[i], xj = d.x0[j], uj = d.u[j];
        a += (xi.x + ui.x) * (xj.y + uj.y) - (xj.x + uj.x) * (xi.y + ui.y);
    }
    return 0.5 * a;
}

template <bool KIN>
__device__ __forceinline__ nfd::PointState load_point(const Dev& d, int p) {
    nfd::PointState q;
    q.s0 = d.sig0[p]; q.s1 = d.sig1[p]; q.s2 = d.sig2[p]; q.s3 = d.sig3[p];
    if (KIN) {
        q.b0 = d.back0[p]; q.b1 = d.back1[p]; q.b2 = d.back2[p]; q.b3 = d.back3[p];
    } else {
        q.b0 = q.b1 = q.b2 = q.b3 = 0.0;
    }
    q.epsP = d.epsP[p];
    q.work = d.work[p];
    return q;
}

template <bool KIN>
__device__ __forceinline__ void store_point(const Dev& d, int p, const nfd::PointState& q) {
    d.sig0[p] = q.s0; d.sig1[p] = q.s1; d.sig2[p] = q.s2; d.sig3[p] = q.s3;
    if (KIN) {
        d.back0[p] = q.b0; d.back1[p] = q.b1; d.back2[p] = q.b2; d.back3[p] = q.b3;
    }
    d.epsP[p] = q.epsP;
    d.work[p] = q.work;
}

__device__ __forceinline__ void load_damage(const Dev& d, int p, nfd::DamagePoint& dp) {
    dp.strength = d.dStrength[p];
    dp.softening = d.dSoftening[p];
    dp.width = d.dWidth[p];
    dp.failureStrain = d.dFailure[p];
    dp.fractureEnergy = d.dGf[p];
    dp.youngs = d.dYoungs[p];
    const double* g = d.dGeom + 5 * (size_t)p;
    dp.area = g[0]; dp.uR = g[1]; dp.uZ = g[2]; dp.vR = g[3]; dp.vZ = g[4];
}

// Strain energy of one point, solver/ExplicitSolver.java:1666-1675 (without the volume).
__device__ __forceinline__ double point_energy_density(const Dev& d, int p, const Mat& m) {
    const double s0 = d.sig0[p], s1 = d.sig1[p], s2 = d.sig2[p], s3 = d.sig3[p];
    const double press = (s0 + s1 + s2) / 3.0;
    const double a = s0 - press, bb = s1 - press, cc = s2 - press;
    return (a * a + bb * bb + cc * cc + 2.0 * s3 * s3) / (4.0 * m.mu) + press * press / (2.0 * m.bulk);
}

// Gauss-point volume on the reference mesh, solver/ExplicitSolver.java:1939-1965.
__device__ __forceinline__ double gaussVolume(const Dev& d, int4 n, int g) {
    const double2 p0 = d.x0[n.x], p1 = d.x0[n.y], p2 = d.x0[n.z], p3 = d.x0[n.w];
    const double G = 0.5773502691896258;
    const double xi = d.ppe == 1 ? 0.0 : ((g == 0 || g == 3) ? -G : G);
    const double eta = d.ppe == 1 ? 0.0 : (g < 2 ? -G : G);
    const double weight = d.ppe == 1 ? 4.0 : 1.0;
    const double xm = 1.0 - xi, xp = 1.0 + xi, em = 1.0 - eta, ep = 1.0 + eta;
    const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em, N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
    const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
    const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
    const double J00 = dN0x * p0.x + dN1x * p1.x + dN2x * p2.x + dN3x * p3.x;
    const double J01 = dN0x * p0.y + dN1x * p1.y + dN2x * p2.y + dN3x * p3.y;
    const double J10 = dN0e * p0.x + dN1e * p1.x + dN2e * p2.x + dN3e * p3.x;
    const double J11 = dN0e * p0.y + dN1e * p1.y + dN2e * p2.y + dN3e * p3.y;
    const double det = J00 * J11 - J01 * J10;
    const double rg = N0 * p0.x + N1 * p1.x + N2 * p2.x + N3 * p3.x;
    return weight * (d.axi ? 2.0 * NFD_PI * rg * det : d.thickness * det);
}

// Erosion test of one element (solver/ExplicitSolver.java:663-678), fused ahead of its own
// assembly: the test reads only the element's own points and the step's new displacements,
// exactly what the reference's separate pass reads. Returns true if the element is gone.
__device__ bool erode_element(const Dev& d, int e, int4 n, double& energy, int& eroded,
                              int& inverted, bool& stale) {
    if (d.gone[e]) return true;
    bool failed = false;
    if (d.dStrength) {
        double sum = 0.0;
        for (int g = 0; g < d.ppe; g++) {
            const int p = e * d.ppe + g;
            sum += nfd::damage_at(d.dStrength[p], d.dSoftening[p], d.dFailure[p], d.epsP[p]);
        }
        failed = sum / d.ppe >= d.erosionDamage;
    }
    const bool inv = currentArea(d, n) <= 0.0;
    if (!failed && !inv) return false;
    d.gone[e] = 1;
    if (d.contact) d.cDead[e] = 1;  // contact.forget(e) :676
    eroded++;
    if (inv && !failed) inverted++;
    const Mat m = load_mat(d, e);
    for (int g = 0; g < d.ppe; g++) energy += point_energy_density(d, e * d.ppe + g, m) * gaussVolume(d, n, g);
    stale = true;
    return true;
}

__device__ __forceinline__ void write_zero_slots(const Dev& d, int e) {
    const double2 z = make_double2(0.0, 0.0);
    d.slot[4 * e] = z;
    d.slot[4 * e + 1] = z;
    d.slot[4 * e + 2] = z;
    d.slot[4 * e + 3] = z;
}

// One element under reduced integration, solver/ExplicitSolver.java:1373-1504.
// REC (the adjoint's recording pass only): *dGammaOut and *status receive the return map's
// multiplier and outcome (j2_update).
// The places of element e's four corner slots in the element-major d.slot.
__device__ __forceinline__ int4 elementSlots(int e) { return make_int4(4 * e, 4 * e + 1, 4 * e + 2, 4 * e + 3); }

// The reference positions, displacements and velocities of an element's four corners.
struct Corners {
    double2 p0, p1, p2, p3, u0, u1, u2, u3, v0, v1, v2, v3;
};
__device__ __forceinline__ Corners load_corners(const Dev& d, int4 n) {
    return Corners{d.x0[n.x], d.x0[n.y], d.x0[n.z], d.x0[n.w], d.u[n.x], d.u[n.y], d.u[n.z], d.u[n.w],
                   d.v[n.x], d.v[n.y], d.v[n.z], d.v[n.w]};
}

// The element on corner values already in registers (c), writing its four force slots to
// slotOut[dest.x .. dest.w] (d.slot at elementSlots(e), or a fused run's node-major slots at
// orderDest).
template <bool DAMAGE, bool KIN, bool REC = false>
__device__ __forceinline__ double reduced_element_at(const Dev& d, int e, const Corners& c, double2* slotOut,
                                       int4 dest, double strainDt, const nfm::Tables& T, double* dGammaOut = nullptr,
                                       int* status = nullptr, int* iterations = nullptr) {
    const Mat fl = load_mat(d, e);
    const double mu = fl.mu;
    const double mid = -0.5 * strainDt;
    const double2 p0 = c.p0, p1 = c.p1, p2 = c.p2, p3 = c.p3;
    const double2 u0 = c.u0, u1 = c.u1, u2 = c.u2, u3 = c.u3;
    const double2 v0 = c.v0, v1 = c.v1, v2 = c.v2, v3 = c.v3;
    double r0, r1, r2, r3, z0, z1, z2, z3;
    if (d.finite) {  // cr / cz :1176-1182
        r0 = p0.x + u0.x + mid * v0.x; r1 = p1.x + u1.x + mid * v1.x;
        r2 = p2.x + u2.x + mid * v2.x; r3 = p3.x + u3.x + mid * v3.x;
        z0 = p0.y + u0.y + mid * v0.y; z1 = p1.y + u1.y + mid * v1.y;
        z2 = p2.y + u2.y + mid * v2.y; z3 = p3.y + u3.y + mid * v3.y;
    } else {
        r0 = p0.x; r1 = p1.x; r2 = p2.x; r3 = p3.x;
        z0 = p0.y; z1 = p1.y; z2 = p2.y; z3 = p3.y;
    }
    const double a0 = v0.x * strainDt, a1 = v1.x * strainDt, a2 = v2.x * strainDt, a3 = v3.x * strainDt;
    const double c0 = v0.y * strainDt, c1 = v1.y * strainDt, c2 = v2.y * strainDt, c3 = v3.y * strainDt;
    double minEdge = NFD_MAX_VALUE;
    if (d.finite) minEdge = nfd::shortestEdge(minEdge, r0, r1, r2, r3, z0, z1, z2, z3);

    const double J00 = 0.25 * (-r0 + r1 + r2 - r3);
    const double J01 = 0.25 * (-z0 + z1 + z2 - z3);
    const double J10 = 0.25 * (-r0 - r1 + r2 + r3);
    const double J11 = 0.25 * (-z0 - z1 + z2 + z3);
    const double det = J00 * J11 - J01 * J10;
    const double inv = 1.0 / det;
    const double d0r = 0.25 * (J01 - J11) * inv;
    const double d1r = 0.25 * (J01 + J11) * inv;
    const double d2r = -d0r;
    const double d3r = -d1r;
    const double d0z = 0.25 * (J10 - J00) * inv;
    const double d1z = -0.25 * (J10 + J00) * inv;
    const double d2z = -d0z;
    const double d3z = -d1z;
    const double dEpsR = d0r * a0 + d1r * a1 + d2r * a2 + d3r * a3;
    const double dEpsZ = d0z * c0 + d1z * c1 + d2z * c2 + d3z * c3;
    const double dRdZ = d0z * a0 + d1z * a1 + d2z * a2 + d3z * a3;
    const double dZdR = d0r * c0 + d1r * c1 + d2r * c2 + d3r * c3;
    const double dGamRZ = dRdZ + dZdR;
    double rg = 0.0, dEpsT = 0.0;
    if (d.axi) {
        rg = 0.25 * (r0 + r1 + r2 + r3);
        dEpsT = 0.25 * (a0 + a1 + a2 + a3) / rg;
    }

    nfd::PointState q = load_point<KIN>(d, e);
    if (d.finite) {
        const double spinDt = 0.5 * (dZdR - dRdZ);
        if (spinDt != 0.0) {
            const double cs = nfd::corot_cos(spinDt), sn = nfd::corot_sin(spinDt);
            nfd::corot_rotate(q.s0, q.s1, q.s3, cs, sn);
            if (KIN) nfd::corot_rotate(q.b0, q.b1, q.b3, cs, sn);
        }
    }
    nfd::DamagePoint dp;
    if (DAMAGE) load_damage(d, e, dp);
    const double dGamma = nfd::j2_update<DAMAGE, KIN, REC>(q, dEpsR, dEpsZ, dEpsT, dGamRZ, fl, strainDt, &dp, T, status, iterations);
    if (REC) *dGammaOut = dGamma;
    store_point<KIN>(d, e, q);
    if (DAMAGE) {
        d.dStrength[e] = dp.strength;
        d.dSoftening[e] = dp.softening;
        d.dWidth[e] = dp.width;
    }

    const double sR = q.s0, sZ = q.s1, sT = q.s2, sRZ = q.s3;
    const double dV = 4.0 * (d.axi ? 2.0 * NFD_PI * rg * det : d.thickness * det);
    double g0r, g1r, g2r, g3r;
    if (d.axi) {
        const double h = sT / rg * dV * 0.25;
        g0r = -((d0r * sR + d0z * sRZ) * dV + h);
        g1r = -((d1r * sR + d1z * sRZ) * dV + h);
        g2r = -((d2r * sR + d2z * sRZ) * dV + h);
        g3r = -((d3r * sR + d3z * sRZ) * dV + h);
    } else {
        g0r = -(d0r * sR + d0z * sRZ) * dV;
        g1r = -(d1r * sR + d1z * sRZ) * dV;
        g2r = -(d2r * sR + d2z * sRZ) * dV;
        g3r = -(d3r * sR + d3z * sRZ) * dV;
    }
    double g0z = -(d0z * sZ + d0r * sRZ) * dV;
    double g1z = -(d1z * sZ + d1r * sRZ) * dV;
    double g2z = -(d2z * sZ + d2r * sRZ) * dV;
    double g3z = -(d3z * sZ + d3r * sRZ) * dV;
    if (d.qhg > 0.0) {
        const double hr = r0 - r1 + r2 - r3;
        const double hz = z0 - z1 + z2 - z3;
        const double y0 = 1.0 - hr * d0r - hz * d0z;
        const double y1 = -1.0 - hr * d1r - hz * d1z;
        const double y2 = 1.0 - hr * d2r - hz * d2z;
        const double y3 = -1.0 - hr * d3r - hz * d3z;
        const double k = d.qhg * mu * dV *
                         (d0r * d0r + d1r * d1r + d2r * d2r + d3r * d3r + d0z * d0z + d1z * d1z +
                          d2z * d2z + d3z * d3z);
        const double qr = d.hgR[e] + k * (y0 * a0 + y1 * a1 + y2 * a2 + y3 * a3);
        const double qz = d.hgZ[e] + k * (y0 * c0 + y1 * c1 + y2 * c2 + y3 * c3);
        d.hgR[e] = qr;
        d.hgZ[e] = qz;
        g0r -= y0 * qr; g1r -= y1 * qr; g2r -= y2 * qr; g3r -= y3 * qr;
        g0z -= y0 * qz; g1z -= y1 * qz; g2z -= y2 * qz; g3z -= y3 * qz;
        d.hgK[e] = k;
    }
    slotOut[dest.x] = make_double2(g0r, g0z);
    slotOut[dest.y] = make_double2(g1r, g1z);
    slotOut[dest.z] = make_double2(g2r, g2z);
    slotOut[dest.w] = make_double2(g3r, g3z);
    return minEdge;
}

template <bool DAMAGE, bool KIN, bool REC = false>
__device__ __forceinline__ double reduced_element(const Dev& d, int e, int4 n, double strainDt,
                                       const nfm::Tables& T, double* dGammaOut = nullptr,
                                       int* status = nullptr, int* iterations = nullptr) {
    return reduced_element_at<DAMAGE, KIN, REC>(d, e, load_corners(d, n), d.slot, elementSlots(e), strainDt, T, dGammaOut,
                                                status, iterations);
}

// One element under full 2x2 integration, solver/ExplicitSolver.java:1219-1339, on corner
// values in registers.
template <bool DAMAGE, bool KIN>
__device__ __forceinline__ double full_element_at(const Dev& d, int e, const Corners& c, double2* slotOut,
                                       int4 dest, double strainDt, const nfm::Tables& T, int* iterations = nullptr) {
    const Mat fl = load_mat(d, e);
    const double mid = -0.5 * strainDt;
    const double2 p0 = c.p0, p1 = c.p1, p2 = c.p2, p3 = c.p3;
    const double2 u0 = c.u0, u1 = c.u1, u2 = c.u2, u3 = c.u3;
    const double2 v0 = c.v0, v1 = c.v1, v2 = c.v2, v3 = c.v3;
    double r0, r1, r2, r3, z0, z1, z2, z3;
    if (d.finite) {
        r0 = p0.x + u0.x + mid * v0.x; r1 = p1.x + u1.x + mid * v1.x;
        r2 = p2.x + u2.x + mid * v2.x; r3 = p3.x + u3.x + mid * v3.x;
        z0 = p0.y + u0.y + mid * v0.y; z1 = p1.y + u1.y + mid * v1.y;
        z2 = p2.y + u2.y + mid * v2.y; z3 = p3.y + u3.y + mid * v3.y;
    } else {
        r0 = p0.x; r1 = p1.x; r2 = p2.x; r3 = p3.x;
        z0 = p0.y; z1 = p1.y; z2 = p2.y; z3 = p3.y;
    }
    const double a0 = v0.x * strainDt, a1 = v1.x * strainDt, a2 = v2.x * strainDt, a3 = v3.x * strainDt;
    const double c0 = v0.y * strainDt, c1 = v1.y * strainDt, c2 = v2.y * strainDt, c3 = v3.y * strainDt;
    double minEdge = NFD_MAX_VALUE;
    if (d.finite) minEdge = nfd::shortestEdge(minEdge, r0, r1, r2, r3, z0, z1, z2, z3);

    const double G = 0.5773502691896258;
    double g0r = 0, g1r = 0, g2r = 0, g3r = 0;
    double g0z = 0, g1z = 0, g2z = 0, g3z = 0;
    for (int g = 0; g < 4; g++) {
        const double xi = (g == 0 || g == 3) ? -G : G;
        const double eta = g < 2 ? -G : G;
        const double xm = 1.0 - xi, xp = 1.0 + xi;
        const double em = 1.0 - eta, ep = 1.0 + eta;
        const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em;
        const double N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
        const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
        const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
        const double J00 = dN0x * r0 + dN1x * r1 + dN2x * r2 + dN3x * r3;
        const double J01 = dN0x * z0 + dN1x * z1 + dN2x * z2 + dN3x * z3;
        const double J10 = dN0e * r0 + dN1e * r1 + dN2e * r2 + dN3e * r3;
        const double J11 = dN0e * z0 + dN1e * z1 + dN2e * z2 + dN3e * z3;
        const double det = J00 * J11 - J01 * J10;
        const double inv = 1.0 / det;
        const double d0r = (J11 * dN0x - J01 * dN0e) * inv;
        const double d1r = (J11 * dN1x - J01 * dN1e) * inv;
        const double d2r = (J11 * dN2x - J01 * dN2e) * inv;
        const double d3r = (J11 * dN3x - J01 * dN3e) * inv;
        const double d0z = (-J10 * dN0x + J00 * dN0e) * inv;
        const double d1z = (-J10 * dN1x + J00 * dN1e) * inv;
        const double d2z = (-J10 * dN2x + J00 * dN2e) * inv;
        const double d3z = (-J10 * dN3x + J00 * dN3e) * inv;
        const double dEpsR = d0r * a0 + d1r * a1 + d2r * a2 + d3r * a3;
        const double dEpsZ = d0z * c0 + d1z * c1 + d2z * c2 + d3z * c3;
        const double dRdZ = d0z * a0 + d1z * a1 + d2z * a2 + d3z * a3;
        const double dZdR = d0r * c0 + d1r * c1 + d2r * c2 + d3r * c3;
        const double dGamRZ = dRdZ + dZdR;
        double rg = 0.0, dEpsT = 0.0;
        if (d.axi) {
            rg = N0 * r0 + N1 * r1 + N2 * r2 + N3 * r3;
            dEpsT = (N0 * a0 + N1 * a1 + N2 * a2 + N3 * a3) / rg;
        }
        const int p = e * 4 + g;
        nfd::PointState q = load_point<KIN>(d, p);
        if (d.finite) {
            const double spinDt = 0.5 * (dZdR - dRdZ);
            if (spinDt != 0.0) {
                const double cs = nfd::corot_cos(spinDt), sn = nfd::corot_sin(spinDt);
                nfd::corot_rotate(q.s0, q.s1, q.s3, cs, sn);
                if (KIN) nfd::corot_rotate(q.b0, q.b1, q.b3, cs, sn);
            }
        }
        nfd::DamagePoint dp;
        if (DAMAGE) load_damage(d, p, dp);
        nfd::j2_update<DAMAGE, KIN>(q, dEpsR, dEpsZ, dEpsT, dGamRZ, fl, strainDt, &dp, T, nullptr, iterations);
        store_point<KIN>(d, p, q);
        if (DAMAGE) {
            d.dStrength[p] = dp.strength;
            d.dSoftening[p] = dp.softening;
            d.dWidth[p] = dp.width;
        }
        const double sR = q.s0, sZ = q.s1, sT = q.s2, sRZ = q.s3;
        const double dV = d.axi ? 2.0 * NFD_PI * rg * det : d.thickness * det;
        if (d.axi) {
            const double h = sT / rg * dV;
            g0r -= (d0r * sR + d0z * sRZ) * dV + N0 * h;
            g1r -= (d1r * sR + d1z * sRZ) * dV + N1 * h;
            g2r -= (d2r * sR + d2z * sRZ) * dV + N2 * h;
            g3r -= (d3r * sR + d3z * sRZ) * dV + N3 * h;
        } else {
            g0r -= (d0r * sR + d0z * sRZ) * dV;
            g1r -= (d1r * sR + d1z * sRZ) * dV;
            g2r -= (d2r * sR + d2z * sRZ) * dV;
            g3r -= (d3r * sR + d3z * sRZ) * dV;
        }
        g0z -= (d0z * sZ + d0r * sRZ) * dV;
        g1z -= (d1z * sZ + d1r * sRZ) * dV;
        g2z -= (d2z * sZ + d2r * sRZ) * dV;
        g3z -= (d3z * sZ + d3r * sRZ) * dV;
    }
    slotOut[dest.x] = make_double2(g0r, g0z);
    slotOut[dest.y] = make_double2(g1r, g1z);
    slotOut[dest.z] = make_double2(g2r, g2z);
    slotOut[dest.w] = make_double2(g3r, g3z);
    return minEdge;
}

template <bool DAMAGE, bool KIN>
__device__ __forceinline__ double full_element(const Dev& d, int e, int4 n, double strainDt,
                                       const nfm::Tables& T, int* iterations = nullptr) {
    return full_element_at<DAMAGE, KIN>(d, e, load_corners(d, n), d.slot, elementSlots(e), strainDt, T, iterations);
}

// Pressure edge k, solver/ExplicitSolver.java:1511-1556 (coordinates at the current, not the
// mid-step, configuration: cr(node, 0.0)).
__device__ void pressure_edge(const Dev& d, int k, double p) {
    const int2 ed = d.pedges[k];
    const size_t s = 4 * ((size_t)d.E + k);
    const double2 xa = d.x0[ed.x], xb = d.x0[ed.y];
    double ra, rb, za, zb;
    if (d.finite) {
        const double2 ua = d.u[ed.x], ub = d.u[ed.y], va = d.v[ed.x], vb = d.v[ed.y];
        ra = xa.x + ua.x + 0.0 * va.x;
        rb = xb.x + ub.x + 0.0 * vb.x;
        za = xa.y + ua.y + 0.0 * va.y;
        zb = xb.y + ub.y + 0.0 * vb.y;
    } else {
        ra = xa.x; rb = xb.x; za = xa.y; zb = xb.y;
    }
    const double tr = rb - ra;
    const double tz = zb - za;
    const double len = __builtin_sqrt(tr * tr + tz * tz);
    if (len == 0.0) {
        d.slot[s] = make_double2(0.0, 0.0);
        d.slot[s + 1] = make_double2(0.0, 0.0);
        return;
    }
    const double nr = tz / len;
    const double nz = -tr / len;
    double wa, wb;
    if (d.axi) {
        const double circumference = 2.0 * NFD_PI * len;
        wa = circumference * (ra / 3.0 + rb / 6.0);
        wb = circumference * (ra / 6.0 + rb / 3.0);
    } else {
        wa = d.thickness * len * 0.5;
        wb = wa;
    }
    d.slot[s] = make_double2(p * nr * wa, p * nz * wa);
    d.slot[s + 1] = make_double2(p * nr * wb, p * nz * wb);
}

template <bool FULL, bool DAMAGE, bool KIN>
// nodeMajor: the first element pass of a fused run writes its slots node-major (slotN at
// orderDest), where the fused step kernel that follows gathers them.
__global__ void __launch_bounds__(BLOCK, (FULL || DAMAGE) ? 2 : 3) k_element(Dev d, int q, int mode, int nodeMajor) {
    Ctl* c = d.ctl;
    Step& st = c->s[q];
    if (mode == MODE_STEP && !st.go) return;
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    // Each element's update reads and writes only that element's own state, so which thread
    // takes which element changes no result. With a schedule, the first 256 patchBlocks
    // threads take patches and the pressure edges follow; erosion's energy is summed per
    // patch (lanes in the patch's own order), so its sum does not depend on the schedule
    // either. The element and its corners are read before the table staging's barrier, so
    // that the two overlap.
    const bool scheduled = d.order != nullptr;
    int e = t, patch = -1, edgeIndex = t - d.E;
    int4 n = make_int4(0, 0, 0, 0), dest = make_int4(0, 0, 0, 0);
    if (scheduled) {
        const int elementThreads = BLOCK * d.patchBlocks;
        edgeIndex = t - elementThreads;
        e = -1;
        if (t < elementThreads) {
            patch = d.patchOf[(size_t)q * 4 * d.patchBlocks + __builtin_amdgcn_readfirstlane(t >> 6)];  // one scalar load a wave
            if (patch >= 0) {
                e = d.order[64 * patch + (t & 63)];
                n = d.orderConn[64 * patch + (t & 63)];
                if (nodeMajor) dest = d.orderDest[64 * patch + (t & 63)];
            }
        }
    } else if (t < d.E) {
        n = d.conn[t];
    }
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    const double strainDt = mode == MODE_INITIAL ? 0.0 : st.dt;
    double edge = NFD_MAX_VALUE;
    bool sawNaN = false;
    double eEnergy = 0.0;
    int eCount = 0, iCount = 0;
    bool stale = false;
    int iterations = 0;
    if (e >= 0 && e < d.E) {
        bool gone = false;
        if (d.gone) {
            gone = mode == MODE_STEP ? erode_element(d, e, n, eEnergy, eCount, iCount, stale) : d.gone[e] != 0;
        }
        if (gone) {
            write_zero_slots(d, e);
        } else {
            if (nodeMajor)
                edge = FULL ? full_element_at<DAMAGE, KIN>(d, e, load_corners(d, n), d.slotN, dest, strainDt, T, &iterations)
                            : reduced_element_at<DAMAGE, KIN>(d, e, load_corners(d, n), d.slotN, dest, strainDt, T, nullptr,
                                                              nullptr, &iterations);
            else
                edge = FULL ? full_element<DAMAGE, KIN>(d, e, n, strainDt, T, &iterations)
                            : reduced_element<DAMAGE, KIN>(d, e, n, strainDt, T, nullptr, nullptr, &iterations);
            sawNaN = edge != edge;
        }
    } else if (edgeIndex >= 0 && edgeIndex < d.edges) {
        pressure_edge(d, edgeIndex, pressureAt(d, st.time));
    }
    if (patch >= 0) {
        // The patch's cost for the next schedule: its slowest lane's Newton iterations.
        int m = iterations;
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        if ((threadIdx.x & 63) == 0) d.patchCost[(size_t)q * d.patches + patch] = (uint8_t)(m < 255 ? m : 255);
    }

    // Shortest edge: exact minimum over non-negative doubles by their bit patterns.
    unsigned long long bits = sawNaN ? MAX_BITS : (unsigned long long)nfm::bits_of(edge);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o < bits ? o : bits;
    }
    const unsigned long long anyNaN = __ballot(sawNaN);
    __shared__ unsigned long long ldsMin[BLOCK / 64];
    __shared__ int ldsNaN;
    if (threadIdx.x == 0) ldsNaN = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        ldsMin[threadIdx.x >> 6] = bits;
        if (anyNaN) ldsNaN = 1;
    }
    if (d.gone && mode == MODE_STEP) {
        __shared__ double ldsE[4 * 1];
        double ev[1] = {eEnergy};
        if (scheduled) {
            // per patch: a fixed shuffle tree over the patch's lanes
            for (int off = 32; off > 0; off >>= 1) ev[0] += __shfl_xor(ev[0], off, 64);
            if ((threadIdx.x & 63) == 0 && patch >= 0) d.erodePartial[patch] = ev[0];
        } else {
            block_sum<1>(ev, ldsE);
            if (threadIdx.x == 0) d.erodePartial[blockIdx.x] = ev[0];
        }
        int a = eCount, b = iCount, sflag = stale ? 1 : 0;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
            sflag |= __shfl_xor(sflag, off, 64);
        }
        if ((threadIdx.x & 63) == 0 && (a | b | sflag)) {
            atomicAdd(&c->erodedCount, a);
            atomicAdd(&c->invertedCount, b);
            if (sflag && d.contact) {
                c->cNAct = 0;  // the rebuild that follows appends from zero
                c->cNSurf = 0;
            }
            if (sflag) c->surfaceStale = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = ldsMin[0];
        for (int w = 1; w < BLOCK / 64; w++) m = ldsMin[w] < m ? ldsMin[w] : m;
        if (m != MAX_BITS) atomicMin(&st.minEdgeBits, m);
        if (ldsNaN) atomicOr(&st.nanEdge, 1);
    }
    if (mode == MODE_STEP && d.wall && blockIdx.x == 0)
        reduce_wall_partials(d, d.wallPartial, d.wallCountPartial, d.nodalBlocks);
}

// ------------------------------------------------------------------ the fused step

// One launch per step where the step has no penalty contact, erosion or pressure edges: the
// node launch that closes step n - 1 and opens step n (k_gn) and the element launch of step n
// (k_element) in one kernel. The host enqueues it in place of the pair k_gn, k_element, so a
// run call is: k_step (opening its first step) ... k_step, then the closing k_gn (at step 0:
// k_element, k_gn, k_element, k_step ... k_step, k_gn).
//
// Each wave takes a patch of 64 elements from the element pass's schedule and first does the
// node launch's work for every node of the patch: the fixed-order gather of step n - 1's force
// slots plus gravity, the CFL step and end test (the same values in every workgroup), and the
// central-difference update with the rigid wall that opens step n. A node on the patch's border
// is also computed by the neighbouring patches, from the same inputs by the same operations,
// so every copy has the same bits; only the owning patch writes its new state and gathered
// force and counts its wall audits. The updated positions and velocities stay in LDS, and the
// element pass of step n reads its corners there. Because a patch reads the state that the
// owners of its border nodes overwrite in the same launch, node state and force slots are
// double-buffered: the launch reads buffer Step::buf and writes the other.
//
// Control: the launch reads Step q and its lead writes Step q ^ 1 (as k_gn does); the element
// phase belongs to Step q ^ 1, so it takes the schedule patchOf[q ^ 1] and writes
// patchCost[q ^ 1]; the last workgroup remakes patchOf[q] from patchCost[q] (the previous
// launch's costs) for the launch after next. The shortest edge goes through Ctl::fMin: the
// launch with rotation index k3 reads slot k3 (or, first after a plain element launch, Step q),
// lowers slot (k3 + 1) % 3 and resets slot (k3 + 2) % 3. Wall partials go to
// wallPartialF[q ^ 1]; the next launch (or the closing k_gn) adds them up when
// Step::pendingWall says so.
//
// from: where this launch's step comes from. STEP_AFTER_STEP: a k_step launch, whose shortest
// edge is in Ctl::fMin[k3]; STEP_AFTER_ELEMENT: the two-launch element pass, whose shortest edge
// is in the Step. OPEN (its own instantiation, so that the other launches' code is untouched):
// the first launch of a run call, which opens its first step from the closed state a run call
// leaves (the step's dt already taken, the forces gathered in d.f), as the node launch's
// PHASE_RESUME does, with the end test made on the device; it also opens the call itself, as
// k_open_run does: the end time, and a step limit maxSteps past the step count it finds (none
// when maxSteps < 0).
constexpr int STEP_AFTER_STEP = 0, STEP_AFTER_ELEMENT = 1;
template <bool FULL, bool DAMAGE, bool KIN, bool WALL, bool OPEN>
__global__ void __launch_bounds__(BLOCK, (FULL || DAMAGE) ? 2 : 3) k_step(Dev d, int q, int k3, int from,
                                                                          double endTime, long long maxSteps) {
    if (blockIdx.x == gridDim.x - 1) {
        schedule_patches(d, d.patchCost + (size_t)q * d.patches, d.patchOf + (size_t)q * 4 * d.patchBlocks,
                         d.patchRun + (size_t)q * 4 * d.patchBlocks);
        return;
    }
    Ctl* c = d.ctl;
    const Step& in = c->s[q];
    Step& out = c->s[q ^ 1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int4 run = d.patchRun[(size_t)(q ^ 1) * 4 * d.patchBlocks +
                                __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64) + wave)];
    const int patch = run.x;
    const int next3 = k3 == 2 ? 0 : k3 + 1, last3 = k3 == 0 ? 2 : k3 - 1;
    constexpr bool open = OPEN;
    if (!open && !in.go) {  // nothing ran: carry the control over
        if (lead) {
            out = in;
            c->fMin[last3] = MAX_BITS;
            c->fNaN[last3] = 0;
        }
        return;
    }
    double dt = in.dt, dtPrev = in.dtPrev;
    if (d.finite && !open) {
        const bool afterElement = from == STEP_AFTER_ELEMENT;
        const bool nanEdge = afterElement ? in.nanEdge != 0 : c->fNaN[k3] != 0;
        const unsigned long long bits = afterElement ? in.minEdgeBits : c->fMin[k3];
        const double minEdge = nanEdge ? NFD_MAX_VALUE : nfm::from_bits(bits);
        dtPrev = in.dt;
        dt = d.cfl * __builtin_sqrt(minEdge * d.massScale) / d.fastestWave;
    }
    const double stepEnd = OPEN ? endTime : in.endTime;
    const long long stepLimit = OPEN ? (maxSteps > 0 ? in.steps + maxSteps : 0x7FFFFFFFFFFFFFFFLL) : in.stepLimit;
    const bool next = in.time < stepEnd && in.steps < stepLimit;
    const int b = in.buf;
    if (lead) {
        Step o = in;
        if (OPEN) {
            o.endTime = stepEnd;
            o.stepLimit = stepLimit;
            c->cOverflow = 0;
            c->runStart = in.steps;
        }
        o.dt = dt;
        o.dtPrev = dtPrev;
        o.go = next ? 1 : 0;
        if (next) {
            o.time = in.time + dt;
            o.steps = in.steps + 1;
        }
        o.nanEdge = 0;
        o.minEdgeBits = MAX_BITS;
        o.buf = next ? b ^ 1 : b;
        o.pendingWall = (WALL && next) ? 1 : 0;
        out = o;
        c->fMin[last3] = MAX_BITS;
        c->fNaN[last3] = 0;
    }
    const double2* slotIn = b ? d.slotNB : d.slotN;
    double2* slotOut = b ? d.slotN : d.slotNB;
    const double2* uIn = b ? d.uB : d.u;
    const double2* vIn = b ? d.vB : d.v;
    const uint8_t* flagsIn = b ? d.flagsB : d.flags;
    double2* uOut = b ? d.u : d.uB;
    double2* vOut = b ? d.v : d.vB;
    uint8_t* flagsOut = b ? d.flags : d.flagsB;

    // The log/exp tables, one copy per wave: their loads are issued first and written to the
    // wave's LDS copy with the node phase's values, so that no workgroup barrier stands between
    // the node records' loads and the gathers that depend on them.
    __shared__ double sLog[BLOCK / 64][97][3];
    __shared__ double sExp[BLOCK / 64][64][2];
    constexpr int TABLE_DOUBLES = 97 * 3 + 64 * 2, TABLE_LOADS = (TABLE_DOUBLES + 63) / 64;
    double tv[TABLE_LOADS];
#pragma unroll
    for (int k = 0; k < TABLE_LOADS; k++) {
        const int i = lane + 64 * k;
        tv[k] = i < 97 * 3 ? (&NFM_LOG_TBL[0][0])[i] : (i < TABLE_DOUBLES ? (&NFM_EXP_TBL[0][0])[i - 97 * 3] : 0.0);
    }
    const int first = run.y, count = run.z;
    // The element phase's element and its corners' LDS slots, loaded now so that the point
    // state's loads (indexed by the element) issue as soon as the node phase is done.
    int eEl = -1;
    uchar4 eLi = make_uchar4(0, 0, 0, 0);
    int4 eDest = make_int4(0, 0, 0, 0);
    if (patch >= 0) {
        eEl = d.order[64 * patch + lane];
        eLi = d.orderLocal[64 * patch + lane];
        eDest = d.orderDest[64 * patch + lane];
    }
    const bool has0 = lane < count, has1 = lane + 64 < count;
    const int c0 = has0 ? d.pnCode[first + lane] : 0, c1 = has1 ? d.pnCode[first + lane + 64] : 0;
    const nfm::Tables T{sLog[wave], sExp[wave]};

    // Node phase: the patch's nodes, at most two per lane, every load of both issued before
    // either is used.
    __shared__ double2 sX0[BLOCK / 64][STEP_PATCH_NODES];
    __shared__ double2 sU[BLOCK / 64][STEP_PATCH_NODES];
    __shared__ double2 sV[BLOCK / 64][STEP_PATCH_NODES];
    const double dtAvg = 0.5 * (dtPrev + dt);
    const double half = 0.5 * d.damping * dtAvg;
    const double decay = (1.0 - half) / (1.0 + half);
    const double gain = dtAvg / (1.0 + half);
    double acc[3] = {0.0, 0.0, 0.0};
    int contacts = 0, stuck = 0;
    {
        const double2 z2 = make_double2(0.0, 0.0);
        const int k0 = c0 & 0x0fffffff, k1 = c1 & 0x0fffffff;
        // a node's slots: one 64-byte node-major record (entries past its slot count unused);
        // an opening launch takes the forces the run call before gathered instead
        double2 a0[4], a1[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            a0[r] = has0 && !open ? slotIn[4 * (size_t)k0 + r] : z2;
            a1[r] = has1 && !open ? slotIn[4 * (size_t)k1 + r] : z2;
        }
        const double2 f0 = has0 && open ? d.f[k0] : z2, f1 = has1 && open ? d.f[k1] : z2;
        const double m0 = has0 ? d.mass[k0] : 1.0, m1 = has1 ? d.mass[k1] : 1.0;
        const double2 x00 = has0 ? d.x0[k0] : z2, x01 = has1 ? d.x0[k1] : z2;
        double2 uu0 = has0 ? uIn[k0] : z2, uu1 = has1 ? uIn[k1] : z2;
        double2 vv0 = has0 ? vIn[k0] : z2, vv1 = has1 ? vIn[k1] : z2;
        uint8_t fl0 = has0 ? flagsIn[k0] : 0, fl1 = has1 ? flagsIn[k1] : 0;
        // The gather in the reference's order (solver/ExplicitSolver.java:1147-1165): the
        // node's slots in order, then gravity (gather_slots' sum).
        auto gathered = [&](int code, const double2 (&a)[4], double2 fOpen, double m) {
            if (open) return fOpen;
            const int slots = ((code >> 28) & 3) + 1;
            double sr = 0.0, sz = 0.0;
            sr += a[0].x; sz += a[0].y;
            if (slots > 1) { sr += a[1].x; sz += a[1].y; }
            if (slots > 2) { sr += a[2].x; sz += a[2].y; }
            if (slots > 3) { sr += a[3].x; sz += a[3].y; }
            return make_double2(sr + m * d.gravityR, sz + m * d.gravityZ);
        };
        // x0.y is the node's reference axial coordinate, z0.
        auto update = [&](int code, int j, int k, double m, double2 x0, double2 fk, double2 uk, double2 vk, uint8_t fl) {
            const bool owned = code < 0;
            if (owned && !open) d.f[k] = fk;
            if (!next) return;
            double a3[3] = {0.0, 0.0, 0.0};
            int cc = 0, cs = 0;
            nodal_core<WALL>(d, m, x0.y, fl, vk, uk, fk, dt, dtAvg, decay, gain, a3, cc, cs);
            if (owned) {
                uOut[k] = uk;
                vOut[k] = vk;
                flagsOut[k] = fl;
                acc[0] += a3[0];
                acc[1] += a3[1];
                acc[2] += a3[2];
                contacts += cc;
                stuck += cs;
            }
            sX0[wave][j] = x0;
            sU[wave][j] = uk;
            sV[wave][j] = vk;
        };
#pragma unroll
        for (int k = 0; k < TABLE_LOADS; k++) {
            const int i = lane + 64 * k;
            if (i < 97 * 3) (&sLog[wave][0][0])[i] = tv[k];
            else if (i < TABLE_DOUBLES) (&sExp[wave][0][0])[i - 97 * 3] = tv[k];
        }
        if (has0) update(c0, lane, k0, m0, x00, gathered(c0, a0, f0, m0), uu0, vv0, fl0);
        if (has1) update(c1, lane + 64, k1, m1, x01, gathered(c1, a1, f1, m1), uu1, vv1, fl1);
    }
    if (!next) {  // the gather that closes the last step; no step follows
        if (WALL && in.pendingWall && blockIdx.x == 0)
            reduce_wall_partials(d, d.wallPartialF + (size_t)q * 3 * d.patchBlocks,
                                 d.wallCountPartialF + (size_t)q * 2 * d.patchBlocks, d.patchBlocks);
        return;
    }
    // This nodal update's wall audits, summed over the wave now (a fixed shuffle tree) and over
    // the workgroup's waves in wave order after the element phase: block_sum's order, with the
    // workgroup's one barrier at the end.
    __shared__ double ldsW[BLOCK / 64][3];
    __shared__ int ldsWi[BLOCK / 64][2];
    if (WALL) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double x = acc[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            acc[k] = x;
        }
        int a = contacts, bb = stuck;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            bb += __shfl_xor(bb, off, 64);
        }
        if (lane == 0) {
            ldsW[wave][0] = acc[0];
            ldsW[wave][1] = acc[1];
            ldsW[wave][2] = acc[2];
            ldsWi[wave][0] = a;
            ldsWi[wave][1] = bb;
        }
    }
    // The element phase reads only its own wave's nodes: a wave-level fence and barrier publish
    // the LDS writes to the other lanes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // Element phase: step n on the corners in LDS.
    double edge = NFD_MAX_VALUE;
    bool sawNaN = false;
    int iterations = 0;
    if (patch >= 0) {
        const int e = eEl;
        if (e >= 0) {
            const uchar4 li = eLi;
            const Corners cr{sX0[wave][li.x], sX0[wave][li.y], sX0[wave][li.z], sX0[wave][li.w],
                             sU[wave][li.x], sU[wave][li.y], sU[wave][li.z], sU[wave][li.w],
                             sV[wave][li.x], sV[wave][li.y], sV[wave][li.z], sV[wave][li.w]};
            edge = FULL ? full_element_at<DAMAGE, KIN>(d, e, cr, slotOut, eDest, dt, T, &iterations)
                        : reduced_element_at<DAMAGE, KIN>(d, e, cr, slotOut, eDest, dt, T, nullptr, nullptr, &iterations);
            sawNaN = edge != edge;
        }
        int m = iterations;
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        if (lane == 0) d.patchCost[(size_t)(q ^ 1) * d.patches + patch] = (uint8_t)(m < 255 ? m : 255);
    }
    // Shortest edge: exact minimum over non-negative doubles by their bit patterns.
    unsigned long long bits = sawNaN ? MAX_BITS : (unsigned long long)nfm::bits_of(edge);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o < bits ? o : bits;
    }
    const unsigned long long anyNaN = __ballot(sawNaN);
    __shared__ unsigned long long ldsMin[BLOCK / 64];
    __shared__ int ldsNaN[BLOCK / 64];
    if (lane == 0) {
        ldsMin[wave] = bits;
        ldsNaN[wave] = anyNaN ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (WALL) {
            double* pw = d.wallPartialF + (size_t)(q ^ 1) * 3 * d.patchBlocks + 3 * blockIdx.x;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double sum = 0.0;
                for (int w = 0; w < BLOCK / 64; w++) sum += ldsW[w][k];
                pw[k] = sum;
            }
            int* pc = d.wallCountPartialF + (size_t)(q ^ 1) * 2 * d.patchBlocks + 2 * blockIdx.x;
            pc[0] = ldsWi[0][0] + ldsWi[1][0] + ldsWi[2][0] + ldsWi[3][0];
            pc[1] = ldsWi[0][1] + ldsWi[1][1] + ldsWi[2][1] + ldsWi[3][1];
        }
        unsigned long long m = ldsMin[0];
        int nan = ldsNaN[0];
        for (int w = 1; w < BLOCK / 64; w++) {
            m = ldsMin[w] < m ? ldsMin[w] : m;
            nan |= ldsNaN[w];
        }
        if (m != MAX_BITS) atomicMin(&c->fMin[next3], m);
        if (nan) atomicOr(&c->fNaN[next3], 1);
    }
    if (WALL && in.pendingWall && blockIdx.x == 0)
        reduce_wall_partials(d, d.wallPartialF + (size_t)q * 3 * d.patchBlocks,
                             d.wallCountPartialF + (size_t)q * 2 * d.patchBlocks, d.patchBlocks);
}

// Sets the current step control's end time and step limit on the device (no host round trip).
__global__ void k_set_limits(Ctl* c, int q, double endTime, long long stepLimit) {
    c->s[q].endTime = endTime;
    c->s[q].stepLimit = stepLimit;
}

// Opens a run call on the device, with no host round trip: clears the term-overflow flag (it
// reports on one call's steps), records the step count the call starts from, and sets the end
// time and the step limit, maxSteps past that count (none when maxSteps < 0). The opening node
// launch that follows (PHASE_RESUME) makes the end test itself.
__global__ void k_open_run(Ctl* c, int q, double endTime, long long maxSteps) {
    Step& s = c->s[q];
    c->cOverflow = 0;
    c->runStart = s.steps;
    s.endTime = endTime;
    s.stepLimit = maxSteps > 0 ? s.steps + maxSteps : 0x7FFFFFFFFFFFFFFFLL;
}

// A checkpoint of the step loop's state (the adjoint's): every array in one launch, regions
// laid out back to back in `mem`, each rounded up to 16 bytes. save: live -> mem, else back.
struct CkptRegion {
    char* live;
    size_t offset, bytes;
};
constexpr int CKPT_REGIONS = 14;
struct CkptList {
    CkptRegion r[CKPT_REGIONS];
};
__global__ void __launch_bounds__(BLOCK) k_ckpt_copy(CkptList list, char* mem, int save) {
    const size_t stride = (size_t)gridDim.x * BLOCK, t0 = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    for (int k = 0; k < CKPT_REGIONS; k++) {
        const CkptRegion r = list.r[k];
        char* a = save ? mem + r.offset : r.live;
        const char* b = save ? r.live : mem + r.offset;
        const size_t words = r.bytes / 16;
        for (size_t i = t0; i < words; i += stride) ((double2*)a)[i] = ((const double2*)b)[i];
        for (size_t i = words * 16 + t0; i < r.bytes; i += stride) a[i] = b[i];
    }
}

// ------------------------------------------------------------------ film frames

// The element quantities a film frame can carry (enum nf_film_field), at most FILM_MAX.
constexpr int FILM_MAX = 16;
struct FilmFields {
    int n;
    int id[FILM_MAX];
};

// Film.capture (render/Film.java:159-189) for the fields the reference's scenes colour by,
// each restated from its accessor in solver/ExplicitSolver.java with the same summation order:
// out[0, N) deformed r (mesh.r + ur, :161-166), out[N, 2N) deformed z, then one block of E
// values per field.
__global__ void k_film(Dev d, FilmFields ff, double* __restrict__ out) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < d.N) {
        const double2 x = d.x0[i], u = d.u[i];
        out[i] = x.x + u.x;
        out[d.N + i] = x.y + u.y;
    }
    if (i >= d.E) return;
    const int e = i;
    const int4 c = d.conn[e];
    const int nodes[4] = {c.x, c.y, c.z, c.w};
    for (int f = 0; f < ff.n; f++) {
        double v = 0.0;
        const int id = ff.id[f];
        if (id == NF_FILM_PLASTIC_STRAIN) {  // elementPlasticStrain :1897-1901
            double sum = 0.0;
            for (int g = 0; g < d.ppe; g++) sum += d.epsP[e * d.ppe + g];
            v = sum / d.ppe;
        } else if (id == NF_FILM_DAMAGE) {  // elementDamage :1764-1772
            if (d.dStrength) {
                double sum = 0.0;
                for (int g = 0; g < d.ppe; g++) {
                    const int p = e * d.ppe + g;
                    sum += nfd::damage_at(d.dStrength[p], d.dSoftening[p], d.dFailure[p], d.epsP[p]);
                }
                v = sum / d.ppe;
            }
        } else if (id == NF_FILM_SPEED) {  // elementSpeed :895-902
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double2 w = d.v[nodes[k]];
                sum += __builtin_sqrt(w.x * w.x + w.y * w.y);
            }
            v = 0.25 * sum;
        } else if (id == NF_FILM_CENTROID_Z) {  // centroidZ, current configuration :1918-1925
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) sum += d.x0[nodes[k]].y + d.u[nodes[k]].y;
            v = 0.25 * sum;
        } else if (id == NF_FILM_ALIVE) {  // !isEroded :627-629
            v = d.gone && d.gone[e] ? 0.0 : 1.0;
        } else {  // centroidStress :1886-1894, and the scenes' von Mises of it (Impact.java:229-232)
            double s4[4] = {0.0, 0.0, 0.0, 0.0};
            for (int g = 0; g < d.ppe; g++) {
                const int p = e * d.ppe + g;
                s4[0] += d.sig0[p];
                s4[1] += d.sig1[p];
                s4[2] += d.sig2[p];
                s4[3] += d.sig3[p];
            }
            for (int k = 0; k < 4; k++) s4[k] /= d.ppe;
            if (id == NF_FILM_MISES) {
                const double a = s4[0] - s4[1], b = s4[1] - s4[2], cc = s4[2] - s4[0];
                v = __builtin_sqrt(0.5 * (a * a + b * b + cc * cc) + 3.0 * s4[3] * s4[3]);
            } else {
                v = s4[id - NF_FILM_STRESS_RR];
            }
        }
        out[2 * d.N + f * d.E + e] = v;
    }
}

// ------------------------------------------------------------------ setup and small kernels

// Row-sum lumped mass (solver/ExplicitSolver.java:1559-1597) as a per-node gather over the
// CSR, whose element order reproduces the reference's accumulation order (element by element,
// Gauss point by Gauss point).
__global__ void k_lump_mass(Dev d, const double* densityOfElement, double* massOut) {
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= d.N) return;
    const double G = 0.5773502691896258;
    double m = 0.0;
    for (int k = d.slotStart[n]; k < d.slotStart[n + 1]; k++) {
        const int s = d.slotOf[k];
        if (s >= 4 * d.E) continue;  // pressure edges carry no mass
        const int e = s >> 2, corner = s & 3;
        const int4 c4 = d.conn[e];
        const double rho = densityOfElement[e];
        const double2 q0 = d.x0[c4.x], q1 = d.x0[c4.y], q2 = d.x0[c4.z], q3 = d.x0[c4.w];
        for (int g = 0; g < 4; g++) {
            const double xi = (g == 0 || g == 3) ? -G : G;
            const double eta = g < 2 ? -G : G;
            const double xm = 1.0 - xi, xp = 1.0 + xi;
            const double em = 1.0 - eta, ep = 1.0 + eta;
            const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em;
            const double N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
            const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
            const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
            const double J00 = dN0x * q0.x + dN1x * q1.x + dN2x * q2.x + dN3x * q3.x;
            const double J01 = dN0x * q0.y + dN1x * q1.y + dN2x * q2.y + dN3x * q3.y;
            const double J10 = dN0e * q0.x + dN1e * q1.x + dN2e * q2.x + dN3e * q3.x;
            const double J11 = dN0e * q0.y + dN1e * q1.y + dN2e * q2.y + dN3e * q3.y;
            const double det = J00 * J11 - J01 * J10;
            const double rg = N0 * q0.x + N1 * q1.x + N2 * q2.x + N3 * q3.x;
            const double dV = d.axi ? 2.0 * NFD_PI * rg * det : d.thickness * det;
            const double Nc = corner == 0 ? N0 : corner == 1 ? N1 : corner == 2 ? N2 : N3;
            m += rho * Nc * dV;
        }
    }
    massOut[n] = m;
}

__global__ void k_scale_mass(double* mass, int N, double factor) {
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n < N) mass[n] *= factor;
}

__global__ void k_fill2(double2* a, int N, double x, double y) {
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n < N) a[n] = make_double2(x, y);
}

// advance() :1069-1075, the displacement half.
__global__ void k_advance_disp(Dev d, int q) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    const double dt = d.ctl->s[q].dt;
    const uint8_t fl = d.flags[i];
    double2 ui = d.u[i];
    const double2 vi = d.v[i];
    if (!(fl & FIXED_R)) ui.x += vi.x * dt;
    if (!(fl & FIXED_Z)) ui.y += vi.y * dt;
    d.u[i] = ui;
}

// Single material points driven through J2.update (solver/J2.java:163-296), one thread per
// path, for the constitutive self-checks.
template <bool DAMAGE>
__global__ void k_point_path(Mat m, int paths, int steps, const double* inc, const double* dmgIn,
                             double* out) {
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= paths) return;
    nfd::PointState q{};
    nfd::DamagePoint dp{};
    if (DAMAGE) {
        const double* g = dmgIn + 7 * (size_t)p;
        dp.failureStrain = g[0];
        dp.fractureEnergy = g[1];
        dp.youngs = m.youngs;
        dp.area = g[2]; dp.uR = g[3]; dp.uZ = g[4]; dp.vR = g[5]; dp.vZ = g[6];
    }
    for (int k = 0; k < steps; k++) {
        const double* x = inc + 5 * ((size_t)p * steps + k);
        const double dg = nfd::j2_update<DAMAGE, true>(q, x[0], x[1], x[2], x[3], m, x[4], &dp, T);
        double* o = out + 14 * ((size_t)p * steps + k);
        o[0] = q.s0; o[1] = q.s1; o[2] = q.s2; o[3] = q.s3;
        o[4] = q.b0; o[5] = q.b1; o[6] = q.b2; o[7] = q.b3;
        o[8] = q.epsP; o[9] = q.work; o[10] = dg;
        o[11] = dp.strength; o[12] = dp.softening; o[13] = dp.width;
    }
}

// Plain device copy, 16 bytes per thread per iteration, grid-stride: the bandwidth probe and
// the known-size kernel the HBM counters are calibrated against.
__global__ void __launch_bounds__(BLOCK) k_copy(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) b[i] = a[i];
}

// Device self-check of the shared transcendental functions (nf_math.h).
__global__ void k_math_eval(int fn, const double* x, const double* y, double* out, int n) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double r;
    switch (fn) {
        case 0: r = nfm::log(x[i]); break;
        case 1: r = nfm::exp(x[i]); break;
        case 2: r = nfm::pow(x[i], y[i]); break;
        case 3: r = nfm::atan2(x[i], y[i]); break;
        case 4: r = nfm::sin(x[i]); break;
        case 5: r = nfm::cos(x[i]); break;
        case 6: r = nfm::hypot(x[i], y[i]); break;
        case 7: r = __builtin_sqrt(x[i]); break;
        case 8: r = x[i] / y[i]; break;
        case 9: r = nfm::log1p(x[i]); break;
        case 10: r = nfm::pow_ordinary(x[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}); break;
        case 11: r = nfm::log_ordinary(x[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}); break;
        case 12:
        case 13: {
            double lx;
            const double pw = nfm::pow_log_ordinary(x[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, lx);
            r = fn == 12 ? pw : lx;
            break;
        }
        case 14:
        case 15: {  // pow_ordinary(x, y) and log_ordinary(y) side by side
            double pw, lg;
            nfm::pow_and_log_ordinary(x[i], y[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, pw, lg);
            r = fn == 14 ? pw : lg;
            break;
        }
        case 16:
        case 17: {  // pow_ordinary(x, y) and pow_ordinary(y, x) side by side
            double p1, p2;
            nfm::pow_pair_ordinary(x[i], y[i], y[i], x[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, p1, p2);
            r = fn == 16 ? p1 : p2;
            break;
        }
        case 24:
        case 25: {  // the Newton loop's pow(x, y) and log(y) for 0 < y < 1, no range selects
            double pw, lg;
            nfm::pow_and_log_ordinary_y01(x[i], y[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, pw, lg);
            r = fn == 24 ? pw : lg;
            break;
        }
        // the hardware's single-precision log2, exp2 and reciprocal, whose error the Newton
        // solve's proven bisections (nfd::ff_bisects) bound
        case 18: r = (double)__builtin_amdgcn_logf((float)x[i]); break;
        case 19: r = (double)__builtin_amdgcn_exp2f((float)x[i]); break;
        case 20: r = (double)__builtin_amdgcn_rcpf((float)x[i]); break;
        case 21:
        case 22:
        case 23: {  // the Johnson-Cook solve: x[5i..5i+4] = norm, stiffness, plastic strain,
                    // thermal factor, dt; y[0..4] = A, B, n, C, eps0 (the unit-rate form when eps0 = 1)
            const double* a = x + 5 * (size_t)i;
            nfd::Mat m{};
            m.a = y[0]; m.b = y[1]; m.n = y[2]; m.c = y[3]; m.eps0 = y[4];
            const nfm::Tables T{NFM_LOG_TBL, NFM_EXP_TBL};
            int proven = 0;
            if (fn == 22)
                r = m.eps0 == 1.0 ? nfd::j2_solve_loop<true, false, false>(a[0], a[1], a[2], a[3], a[4], m, T)
                                  : nfd::j2_solve_loop<false, false, false>(a[0], a[1], a[2], a[3], a[4], m, T);
            else
                r = m.eps0 == 1.0 ? nfd::j2_solve_loop<true>(a[0], a[1], a[2], a[3], a[4], m, T, nullptr, nullptr, &proven)
                                  : nfd::j2_solve_loop<false>(a[0], a[1], a[2], a[3], a[4], m, T, nullptr, nullptr, &proven);
            if (fn == 23) r = proven;
            break;
        }
        default: r = 0.0;
    }
    out[i] = r;
}

inline int blocks_for(int n) { return (n + BLOCK - 1) / BLOCK; }

#ifndef NF_ELEMENT_ORDER
#define NF_ELEMENT_ORDER 2
#endif

// The element pass's thread order. A wave runs as long as its slowest lane, and in a plastic
// zone the Johnson-Cook Newton solve's iteration count varies smoothly in space, so a wave of
// 64 neighbouring elements in a compact patch wastes fewer lane-iterations than 64 elements in
// a line. Elements are sorted by the cell of their centroid in a square grid of side h (the
// mesh's shortest edge, so a lattice puts one element in each cell): 8 x 8 cells to a wave,
// four such patches side by side (8 x 32 cells) to a workgroup, workgroups row by row. Ties
// (several centroids in one cell of an irregular mesh) keep index order.
static std::vector<int> element_order(int E, const int* conn, const double* r, const double* z, double h) {
    std::vector<int> order(E);
    for (int e = 0; e < E; e++) order[e] = e;
    if (NF_ELEMENT_ORDER == 0 || E == 0 || !(h > 0.0)) return order;
    std::vector<double> rc(E), zc(E);
    double rmin = HUGE_VAL, zmin = HUGE_VAL, rmax = -HUGE_VAL;
    for (int e = 0; e < E; e++) {
        const int* c = conn + 4 * (size_t)e;
        rc[e] = 0.25 * (r[c[0]] + r[c[1]] + r[c[2]] + r[c[3]]);
        zc[e] = 0.25 * (z[c[0]] + z[c[1]] + z[c[2]] + z[c[3]]);
        rmin = std::min(rmin, rc[e]);
        rmax = std::max(rmax, rc[e]);
        zmin = std::min(zmin, zc[e]);
    }
    auto cell = [&](double x) { return (int64_t)std::min(std::floor(x / h), 2147483647.0); };
    const int64_t cols = cell(rmax - rmin) + 1;
    std::vector<uint64_t> key(E);
    for (int e = 0; e < E; e++) {
        const int64_t I = cell(rc[e] - rmin), J = cell(zc[e] - zmin);
        uint64_t k;
        if (NF_ELEMENT_ORDER == 1) {  // Morton (Z-order) of the cell
            k = 0;
            for (int b = 0; b < 32; b++) k |= (uint64_t)((I >> b) & 1) << (2 * b) | (uint64_t)((J >> b) & 1) << (2 * b + 1);
        } else if (NF_ELEMENT_ORDER == 3) {  // 16 x 16 workgroups of 2 x 2 patches
            const int64_t bc = (cols + 15) >> 4;
            k = ((((uint64_t)(J >> 4) * bc + (I >> 4)) * 4 + ((J >> 3) & 1) * 2 + ((I >> 3) & 1)) * 64) + (J & 7) * 8 + (I & 7);
        } else {  // 8 x 32 workgroups of four patches side by side
            const int64_t bc = (cols + 31) >> 5;
            k = ((((uint64_t)(J >> 3) * bc + (I >> 5)) * 4 + ((I >> 3) & 3)) * 64) + (J & 7) * 8 + (I & 7);
        }
        key[e] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    if (NF_ELEMENT_ORDER == 4) {  // the 8 x 32 workgroups, dealt out with a stride
        const int nb = blocks_for(E);
        std::vector<int> tmp(order);
        int stride = 97;
        while (std::gcd(stride, nb) != 1) stride++;
        int w = 0;
        for (int k = 0; k < nb; k++) {
            const int src = (int)(((int64_t)k * stride) % nb);
            for (int t = src * BLOCK; t < std::min(E, (src + 1) * BLOCK); t++) tmp[w++] = order[t];
        }
        order = tmp;
    }
    return order;
}
// Workgroups of a launch that gives each thread `chunk` items.
inline int blocks_for_chunks(int n, int chunk) { return std::max(1, (n + BLOCK * chunk - 1) / (BLOCK * chunk)); }

template <typename T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return NF_OK;
    HIP_TRY(hipMalloc((void**)p, sizeof(T) * count));
    HIP_TRY(hipMemset(*p, 0, sizeof(T) * count));
    return NF_OK;
}

#define TRY(expr)               \
    do {                        \
        int rc_ = (expr);       \
        if (rc_) return rc_;    \
    } while (0)

struct HostMaterial {
    nf_material raw;
    double lambda() const {
        const double e = raw.youngs_modulus, nu = raw.poisson_ratio;
        return e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    }
    double mu() const { return raw.youngs_modulus / (2.0 * (1.0 + raw.poisson_ratio)); }
    double wave() const { return std::sqrt((lambda() + 2.0 * mu()) / raw.density); }
};

// Unrolls one material into the kernels' constants (solver/ExplicitSolver.java:289-299,
// J2.Flow.of solver/J2.java:108-113).
Mat make_mat(const nf_material& raw, int formulation) {
    HostMaterial hm{raw};
    Mat t;
    std::memset(&t, 0, sizeof t);
    const double lam = formulation == NF_PLANE_STRESS
                           ? 2.0 * hm.lambda() * hm.mu() / (hm.lambda() + 2.0 * hm.mu())
                           : hm.lambda();
    t.lambda = lam;
    t.mu = hm.mu();
    t.bulk = lam + 2.0 * hm.mu() / 3.0;
    t.density = raw.density;
    t.sy0 = raw.yield_stress;
    t.hIso = raw.isotropic_hardening;
    t.hKin = raw.kinematic_hardening;
    t.youngs = raw.youngs_modulus;
    t.hasJC = raw.has_johnson_cook;
    if (t.hasJC) {
        const nf_johnson_cook& j = raw.johnson_cook;
        t.a = j.a; t.b = j.b; t.n = j.n; t.c = j.c; t.m = j.m;
        t.eps0 = j.reference_strain_rate; t.troom = j.room_temperature;
        t.tmelt = j.melt_temperature; t.cp = j.specific_heat; t.beta = j.taylor_quinney;
        t.thermalCoefficient = j.taylor_quinney / (raw.density * j.specific_heat);
    }
    return t;
}

int check_material(const nf_material* m) {
    if (!m) return fail(NF_E_ARGUMENT, "material must not be null");
    if (m->youngs_modulus <= 0) return fail(NF_E_ARGUMENT, "E must be positive");
    if (m->poisson_ratio <= -1.0 || m->poisson_ratio >= 0.5)
        return fail(NF_E_ARGUMENT, "nu must lie in (-1, 0.5)");
    if (m->density <= 0) return fail(NF_E_ARGUMENT, "density must be positive");
    if (m->yield_stress <= 0) return fail(NF_E_ARGUMENT, "yield stress must be positive");
    if (m->isotropic_hardening < 0 || m->kinematic_hardening < 0)
        return fail(NF_E_ARGUMENT, "hardening moduli must be non-negative");
    if (m->has_johnson_cook) {
        const nf_johnson_cook& j = m->johnson_cook;
        if (j.a <= 0) return fail(NF_E_ARGUMENT, "A must be positive");
        if (j.b < 0) return fail(NF_E_ARGUMENT, "B must be non-negative");
        if (j.n <= 0 || j.n > 1) return fail(NF_E_ARGUMENT, "n must lie in (0, 1]");
        if (j.c < 0) return fail(NF_E_ARGUMENT, "C must be non-negative");
        if (j.m <= 0) return fail(NF_E_ARGUMENT, "m must be positive");
        if (j.reference_strain_rate <= 0) return fail(NF_E_ARGUMENT, "reference strain rate must be positive");
        if (j.melt_temperature <= j.room_temperature)
            return fail(NF_E_ARGUMENT, "melt temperature must exceed room temperature");
        if (j.specific_heat <= 0) return fail(NF_E_ARGUMENT, "specific heat must be positive");
        if (j.taylor_quinney < 0 || j.taylor_quinney > 1)
            return fail(NF_E_ARGUMENT, "Taylor-Quinney fraction must lie in [0, 1]");
        if (m->yield_stress != j.a) return fail(NF_E_ARGUMENT, "yield stress disagrees with the Johnson-Cook A");
    }
    return NF_OK;
}

}  // namespace

// ------------------------------------------------------------------ the handle

#include "tangent.h"
#include "adjoint.h"

struct nf_solver {
    hipStream_t stream = nullptr;
    Dev d{};
    int q = 0;  // control buffer holding the current state
    Ctl* ctl = nullptr;
    int N = 0, E = 0, P = 0, edges = 0, ppe = 1;
    int formulation = 0, integration = 0, kinematics = 0;
    double thickness = 1.0, cfl = 0.5;
    std::vector<HostMaterial> materials;
    std::vector<int> materialOf;
    std::vector<double> r, z;
    std::vector<int> conn, pedges;
    std::vector<int> slotStart, slotOf;
    std::vector<uint8_t> flagsHost;  // fixed flags mirror (onWall lives on the device)
    double fastestWave = 0.0, dt0 = 0.0, massScale = 1.0;
    bool hasWall = false, hasDamage = false, hasErosion = false, hasContact = false;
    bool anyJC = false;
    bool hasKin = false;  // some material has kinematic hardening
    bool tangent = false;  // parameter sensitivities are carried (nf_enable_sensitivity)
    // Discrete adjoint (nf_adjoint_enable): checkpoints of the whole state every adjK steps
    // of the forward run, then the reverse sweep of nf_adjoint_gradient.
    bool adjoint = false;
    int adjK = 64;
    struct Checkpoint {
        long long steps = 0;
        int q = 0;
        char* mem = nullptr;
    };
    std::vector<Checkpoint> checkpoints;  // after 0, K, 2K, ... steps
    Checkpoint finalState;                // after the last recorded step
    ADev ad{};
    bool adjAllocated = false;
    int cFCapAlloc = 0;    // capture-grid entries allocated (nf_contact_set_limits may use fewer)
    TDev td{};
    double wallPlane = 0.0, wallFriction = 0.0;
    // Optional per-kernel timing with HIP events on the handle's stream (nf_profile).
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<int, std::pair<int, int>>> evPending;  // kernel id, (start, stop)
    double kernelMs[4] = {0.0, 0.0, 0.0, 0.0};
    long long kernelCount[4] = {0, 0, 0, 0};
    // device buffers owned here
    std::vector<void*> owned;
    double* densityOfElement = nullptr;
    double* filmBuf = nullptr;  // film frame staging (nf_film_capture), grown on demand
    size_t filmCap = 0;

    ~nf_solver() {
        if (stream) (void)hipStreamSynchronize(stream);
        dropGraphs();
        if (filmBuf) (void)hipFree(filmBuf);
        for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
        for (hipEvent_t e : statusEvent)
            if (e) (void)hipEventDestroy(e);
        if (statusHost) (void)hipHostFree(statusHost);
        if (ctlHost) (void)hipHostFree(ctlHost);
        for (Checkpoint& c : checkpoints)
            if (c.mem) (void)hipFree(c.mem);
        if (finalState.mem) (void)hipFree(finalState.mem);
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int event(int index, hipEvent_t* out) {
        while ((int)evPool.size() <= index) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            evPool.push_back(e);
        }
        *out = evPool[index];
        return NF_OK;
    }

    // Runs the launch `fn(start, stop)`, with a fresh event pair when profiling.
    // Events are handed out from a counter, so a timed launch sequence may contain another.
    int evUsed = 0;
    template <typename F>
    int timed(int kernel, F fn) {
        if (!profiling) return fn(nullptr, nullptr);
        const int base = evUsed;
        evUsed += 2;
        hipEvent_t a, b;
        TRY(event(base, &a));
        TRY(event(base + 1, &b));
        TRY(fn(a, b));
        evPending.push_back({kernel, {base, base + 1}});
        return NF_OK;
    }

    int harvest() {
        if (evPending.empty()) return NF_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        for (auto& p : evPending) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, evPool[p.second.first], evPool[p.second.second]));
            kernelMs[p.first] += ms;
            kernelCount[p.first] += 1;
        }
        evPending.clear();
        evUsed = 0;
        return NF_OK;
    }

    template <typename T>
    int alloc(T** p, size_t count) {
        TRY(dev_alloc(p, count));
        if (*p) owned.push_back((void*)*p);
        return NF_OK;
    }

    double matE(int e) const {
        return materials[materialOf.empty() ? 0 : materialOf[e]].raw.youngs_modulus;
    }

    int sync() {
        HIP_TRY(hipStreamSynchronize(stream));
        return NF_OK;
    }

    // The control block's host copy goes through pinned memory (ctlHost), which the copy
    // engine reads and writes directly, without a staging copy. A run call ends by reading the
    // control block with the device idle; until the next entry point that may change it
    // (touch()), readCtl returns that copy without waiting on the device, so that a harness
    // asking for the time after every step pays no second round trip.
    Ctl ctlCache;
    bool ctlCached = false;
    void touch() { ctlCached = false; }
    int readCtl(Ctl& out) {
        if (ctlCached) {
            out = ctlCache;
            return NF_OK;
        }
        if (!ctlHost) HIP_TRY(hipHostMalloc((void**)&ctlHost, sizeof(Ctl), hipHostMallocDefault));
        HIP_TRY(hipMemcpyAsync(ctlHost, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        out = *ctlHost;
        return NF_OK;
    }

    int writeCtl(const Ctl& in) {
        fastOpen = false;
        touch();
        HIP_TRY(hipMemcpyAsync(ctl, &in, sizeof(Ctl), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return NF_OK;
    }

    // The step buffer that holds the solver's current state.
    int readStep(Step& out) {
        Ctl h;
        TRY(readCtl(h));
        out = h.s[q];
        return NF_OK;
    }

    int writeStep(const Step& in) {
        fastOpen = false;
        touch();
        HIP_TRY(hipMemcpyAsync(&ctl->s[q], &in, sizeof(Step), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return NF_OK;
    }

    // Launches with optional start/stop events recorded by the dispatch itself
    // (hipExtLaunchKernelGGL), so a profiled duration is the kernel's own, not the gap
    // around it.
    template <bool FULL, bool DAMAGE, bool KIN>
    int launchElementT(int mode, int parity, hipEvent_t a, hipEvent_t b, int nodeMajor) {
        // scheduled: patch workgroups, then the pressure edges (k_element)
        const int grid = d.order ? d.patchBlocks + blocks_for(edges) : blocks_for(E + edges);
        if (a) hipExtLaunchKernelGGL((k_element<FULL, DAMAGE, KIN>), dim3(grid), dim3(BLOCK), 0, stream, a, b, 0, d, parity, mode, nodeMajor);
        else hipLaunchKernelGGL((k_element<FULL, DAMAGE, KIN>), dim3(grid), dim3(BLOCK), 0, stream, d, parity, mode, nodeMajor);
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    // nodeMajor: a fused run's first element pass (node-major slots, k_element).
    int launchElement(int mode, int parity, hipEvent_t a = nullptr, hipEvent_t b = nullptr, int nodeMajor = 0) {
        switch ((integration == NF_FULL ? 4 : 0) | (hasDamage ? 2 : 0) | (hasKin ? 1 : 0)) {
            case 0: return launchElementT<false, false, false>(mode, parity, a, b, nodeMajor);
            case 1: return launchElementT<false, false, true>(mode, parity, a, b, nodeMajor);
            case 2: return launchElementT<false, true, false>(mode, parity, a, b, nodeMajor);
            case 3: return launchElementT<false, true, true>(mode, parity, a, b, nodeMajor);
            case 4: return launchElementT<true, false, false>(mode, parity, a, b, nodeMajor);
            case 5: return launchElementT<true, false, true>(mode, parity, a, b, nodeMajor);
            case 6: return launchElementT<true, true, false>(mode, parity, a, b, nodeMajor);
            default: return launchElementT<true, true, true>(mode, parity, a, b, nodeMajor);
        }
    }

    template <bool WALL>
    int launchGNT(int phase, int parity, int gatherFirst, int part, hipEvent_t a, hipEvent_t b) {
        const int minSlot = gnMinSlot, nodeMajor = gnNodeMajor ? 1 : 0;
        // the gather before the contact launches carries the contact's segment rebuild
        // one workgroup more for the element pass's schedule (k_gn), unless the launch closes
        // a fused run (gnNodeMajor): the next run call opens with k_step, and every schedule
        // buffer always holds a whole schedule, so none is made there. With contact the
        // contact search's extra workgroup makes it (launchContact).
        const int schedule = d.order && part == PART_ALL && !gnNodeMajor ? 1 : 0;
        const int grid = blocks_for(N) + (part == PART_GATHER ? blocks_for_chunks(4 * E, C_CHUNK_SEG) : 0) + schedule;
        if (a) hipExtLaunchKernelGGL(k_gn<WALL>, dim3(grid), dim3(BLOCK), 0, stream, a, b, 0, d, parity, phase, gatherFirst, part, minSlot, nodeMajor, schedule);
        else hipLaunchKernelGGL(k_gn<WALL>, dim3(grid), dim3(BLOCK), 0, stream, d, parity, phase, gatherFirst, part, minSlot, nodeMajor, schedule);
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    int launchGNPart(int phase, int parity, int gatherFirst, int part, hipEvent_t a = nullptr, hipEvent_t b = nullptr) {
        return hasWall ? launchGNT<true>(phase, parity, gatherFirst, part, a, b)
                       : launchGNT<false>(phase, parity, gatherFirst, part, a, b);
    }

    // The contact launches between the gather and the rest of the node pass
    // (advanceMaterialAndAssemble :1100-1108). checkGo: inside a step, which runs only if
    // the device's end test let it.
    int launchContact(int parity, int checkGo, hipEvent_t a = nullptr, hipEvent_t b = nullptr) {
        const int g4 = std::min(blocks_for(4 * E), CONTACT_GRID);
        const int gn = std::min(blocks_for(N), CONTACT_GRID);
        if (a) HIP_TRY(hipEventRecord(a, stream));
        const int nb = blocks_for_chunks(N, C_CHUNK_NODE);
        hipLaunchKernelGGL(k_c_nodes_geometry, dim3(nb + g4), dim3(BLOCK), 0, stream, d, parity, checkGo, nb);
        hipLaunchKernelGGL(k_c_grid, dim3(3 * g4), dim3(BLOCK), 0, stream, d, parity, checkGo);
        // one workgroup more for the element pass's schedule, which the gather before these
        // launches (too short to hide the sort) leaves to the search
        const int schedule = d.order ? 1 : 0;
        hipLaunchKernelGGL(k_c_search, dim3(gn + schedule), dim3(BLOCK), 0, stream, d, parity, checkGo, schedule);
        hipLaunchKernelGGL(k_c_apply, dim3(d.cApplyBlocks), dim3(BLOCK), 0, stream, d, parity, checkGo);
        if (b) HIP_TRY(hipEventRecord(b, stream));
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    // The node-centred launch of an assembly that gathered (phase, gatherFirst), with the
    // contact launches spliced in after the gather when contact is on.
    int launchGN(int phase, int parity, int gatherFirst, hipEvent_t a = nullptr, hipEvent_t b = nullptr) {
        const bool gathers = phase == PHASE_STEP || gatherFirst;
        if (!hasContact || !gathers) return launchGNPart(phase, parity, gatherFirst, PART_ALL, a, b);
        TRY(launchGNPart(phase, parity, gatherFirst, PART_GATHER));
        TRY(timed(2, [&](hipEvent_t x, hipEvent_t y) { return launchContact(parity, phase == PHASE_STEP, x, y); }));
        return launchGNPart(phase, parity, gatherFirst, PART_REST, a, b);
    }

    // One step on the stream: the element pass of the open step, then the launch that
    // gathers it and opens the next. The device decides whether either does anything.
    int launchTangent(int mode, int phase, int gatherFirst) {
        if (mode >= 0) {
            hipLaunchKernelGGL(k_t_element, dim3(blocks_for(2 * E)), dim3(BLOCK), 0, stream, d, td, q, mode);
            if (mode == MODE_STEP) hipLaunchKernelGGL(k_t_argmin, dim3(blocks_for(E)), dim3(BLOCK), 0, stream, d, td, q);
        }
        if (phase >= 0) {
            if (hasWall) hipLaunchKernelGGL(k_t_gn<true>, dim3(blocks_for(2 * N)), dim3(BLOCK), 0, stream, d, td, q, phase, gatherFirst);
            else hipLaunchKernelGGL(k_t_gn<false>, dim3(blocks_for(2 * N)), dim3(BLOCK), 0, stream, d, td, q, phase, gatherFirst);
        }
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    int enqueueStep() {
        if (tangent) {
            TRY(launchTangent(MODE_STEP, PHASE_STEP, 0));
            q ^= 1;
            return NF_OK;
        }
        if (fusedRun) {
            // k_step opening the call's first step (past step 0), or k_element for it (at step
            // 0), then k_step for each later step; closeFused adds the closing k_gn.
            if (openPending) {
                const int k3 = (int)(fusedCount % 3);
                TRY(timed(3, [&](hipEvent_t a, hipEvent_t b) { return launchStep(q, k3, STEP_OPEN, a, b); }));
                openPending = false;
                openElement = true;
                lastK3 = k3;
                fusedCount++;
                q ^= 1;
                return NF_OK;
            }
            if (!openElement) {
                TRY(timed(1, [&](hipEvent_t a, hipEvent_t b) { return launchElement(MODE_STEP, q, a, b, 1); }));
                openElement = true;
                lastK3 = -1;
                return NF_OK;
            }
            const int k3 = (int)(fusedCount % 3);
            TRY(timed(3, [&](hipEvent_t a, hipEvent_t b) {
                return launchStep(q, k3, lastK3 < 0 ? STEP_AFTER_ELEMENT : STEP_AFTER_STEP, a, b);
            }));
            lastK3 = k3;
            fusedCount++;
            q ^= 1;
            return NF_OK;
        }
        TRY(timed(1, [&](hipEvent_t a, hipEvent_t b) { return launchElement(MODE_STEP, q, a, b); }));
        TRY(timed(0, [&](hipEvent_t a, hipEvent_t b) { return launchGN(PHASE_STEP, q, 0, a, b); }));
        q ^= 1;
        return NF_OK;
    }

    // ---- the fused step (k_step): one launch per step without contact, erosion, pressure edges
    bool fusedRun = false;      // this run call enqueues fused steps
    bool openElement = false;   // an element pass is enqueued whose closing k_gn is not
    bool openPending = false;   // the run call's first step is still to be opened (STEP_OPEN)
    double openEnd = 0.0;       // that call's end time and step count (k_step's opening)
    long long openMaxSteps = 0;
    int lastK3 = -1;            // rotation index of the last k_step since that element pass
    long long fusedCount = 0;   // k_step launches so far (their fMin rotation)
    int gnMinSlot = -1;         // k_gn's minSlot argument for the next node launch
    bool gnNodeMajor = false;   // k_gn's nodeMajor argument for the next node launch
    bool fusedDisabled = false; // NEOFIZ_FUSED=0

    bool fusedEligible() const {
        return d.pnStart && !fusedDisabled && !useGraphs && !tangent && !adjoint && !hasContact &&
               !d.gone && edges == 0;
    }

    int fusedBuffers() {
        if (d.uB) return NF_OK;  // node state and node-major slots, both buffers
        TRY(alloc(&d.uB, N));
        TRY(alloc(&d.vB, N));
        TRY(alloc(&d.flagsB, N));
        TRY(alloc(&d.slotN, 4 * (size_t)N));
        TRY(alloc(&d.slotNB, 4 * (size_t)N));
        TRY(alloc(&d.wallPartialF, 2 * 3 * (size_t)d.patchBlocks));
        TRY(alloc(&d.wallCountPartialF, 2 * 2 * (size_t)d.patchBlocks));
        return NF_OK;
    }

    template <bool FULL, bool DAMAGE, bool KIN, bool WALL>
    int launchStepT(int parity, int k3, int from, hipEvent_t a, hipEvent_t b) {
        const int grid = d.patchBlocks + 1;  // the last workgroup remakes the schedule
        const double e = openEnd;
        const long long m = openMaxSteps;
        if (from == STEP_OPEN) {
            if (a) hipExtLaunchKernelGGL((k_step<FULL, DAMAGE, KIN, WALL, true>), dim3(grid), dim3(BLOCK), 0, stream, a, b, 0, d, parity, k3, from, e, m);
            else hipLaunchKernelGGL((k_step<FULL, DAMAGE, KIN, WALL, true>), dim3(grid), dim3(BLOCK), 0, stream, d, parity, k3, from, e, m);
        } else {
            if (a) hipExtLaunchKernelGGL((k_step<FULL, DAMAGE, KIN, WALL, false>), dim3(grid), dim3(BLOCK), 0, stream, a, b, 0, d, parity, k3, from, e, m);
            else hipLaunchKernelGGL((k_step<FULL, DAMAGE, KIN, WALL, false>), dim3(grid), dim3(BLOCK), 0, stream, d, parity, k3, from, e, m);
        }
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    // from: STEP_AFTER_STEP, STEP_AFTER_ELEMENT or STEP_OPEN (the OPEN instantiation)
    static constexpr int STEP_OPEN = 2;
    int launchStep(int parity, int k3, int from, hipEvent_t a, hipEvent_t b) {
        switch ((integration == NF_FULL ? 8 : 0) | (hasDamage ? 4 : 0) | (hasKin ? 2 : 0) | (hasWall ? 1 : 0)) {
            case 0: return launchStepT<false, false, false, false>(parity, k3, from, a, b);
            case 1: return launchStepT<false, false, false, true>(parity, k3, from, a, b);
            case 2: return launchStepT<false, false, true, false>(parity, k3, from, a, b);
            case 3: return launchStepT<false, false, true, true>(parity, k3, from, a, b);
            case 4: return launchStepT<false, true, false, false>(parity, k3, from, a, b);
            case 5: return launchStepT<false, true, false, true>(parity, k3, from, a, b);
            case 6: return launchStepT<false, true, true, false>(parity, k3, from, a, b);
            case 7: return launchStepT<false, true, true, true>(parity, k3, from, a, b);
            case 8: return launchStepT<true, false, false, false>(parity, k3, from, a, b);
            case 9: return launchStepT<true, false, false, true>(parity, k3, from, a, b);
            case 10: return launchStepT<true, false, true, false>(parity, k3, from, a, b);
            case 11: return launchStepT<true, false, true, true>(parity, k3, from, a, b);
            case 12: return launchStepT<true, true, false, false>(parity, k3, from, a, b);
            case 13: return launchStepT<true, true, false, true>(parity, k3, from, a, b);
            case 14: return launchStepT<true, true, true, false>(parity, k3, from, a, b);
            default: return launchStepT<true, true, true, true>(parity, k3, from, a, b);
        }
    }

    // Ends a run of fused steps: the last element pass's gather (k_gn), which also brings the
    // node state back to buffer 0 and adds the last fused launch's wall partials.
    int closeFused() {
        if (!openElement) return NF_OK;
        gnMinSlot = lastK3 >= 0 ? (lastK3 + 1) % 3 : -1;
        gnNodeMajor = true;
        const int rc = timed(0, [&](hipEvent_t a, hipEvent_t b) { return launchGN(PHASE_STEP, q, 0, a, b); });
        gnMinSlot = -1;
        gnNodeMajor = false;
        TRY(rc);
        q ^= 1;
        openElement = false;
        lastK3 = -1;
        return NF_OK;
    }

    // A hipGraph of GRAPH_STEPS steps starting at parity p (an even count, so the parity is
    // the same at both ends). Captured on first use and whenever a setter changes a uniform.
    static constexpr int GRAPH_STEPS = 32;
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    Step* statusHost = nullptr;  // pinned
    Ctl* ctlHost = nullptr;      // pinned (readCtl)
    // The last run call ended past step 0 and nothing has written the control block from the
    // host since: the next call opens on the device (k_open_run), with no host round trip.
    bool fastOpen = false;
    hipEvent_t statusEvent[2] = {nullptr, nullptr};
    bool graphDirty = true;
    bool useGraphs = false;  // NEOFIZ_GRAPHS=1 captures blocks of steps in hipGraphs instead

    void dropGraphs() {
        for (int k = 0; k < 2; k++)
            if (graphExec[k]) {
                (void)hipGraphExecDestroy(graphExec[k]);
                graphExec[k] = nullptr;
            }
        graphDirty = false;
    }

    int captureGraph(int p) {
        hipGraph_t g = nullptr;
        const int q0 = q;
        q = p;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = NF_OK;
        for (int k = 0; k < GRAPH_STEPS && rc == NF_OK; k++) rc = enqueueStep();
        hipError_t err = hipStreamEndCapture(stream, &g);
        q = q0;
        if (rc) return rc;
        if (err != hipSuccess) return fail(NF_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(err));
        err = hipGraphInstantiate(&graphExec[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) return fail(NF_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(err));
        return NF_OK;
    }

    int launchSteps(int n) {
        if (useGraphs && !profiling && n == GRAPH_STEPS) {
            if (graphDirty) dropGraphs();
            if (!graphExec[q]) TRY(captureGraph(q));
            HIP_TRY(hipGraphLaunch(graphExec[q], stream));
            return NF_OK;  // GRAPH_STEPS is even: the parity is unchanged
        }
        for (int k = 0; k < n; k++) TRY(enqueueStep());
        return NF_OK;
    }

    // advanceMaterialAndAssemble(strainDt) :1081-1115 outside a step: the initial
    // zero-interval assembly is followed by the first nodal update (PHASE_START with
    // gatherFirst); advance() ends with a bare gather.
    int assembleAdvance() {
        if (tangent || adjoint) return fail(NF_E_STATE, "advance() is not differentiated; sensitivities follow step() only");
        TRY(launchElement(MODE_ADVANCE, q));
        TRY(launchGN(PHASE_ASSEMBLE_END, q, 1));
        q ^= 1;
        return NF_OK;
    }

    // ---------------------------------------------------------------- discrete adjoint
    // A checkpoint holds every array the step loop reads or writes across steps, and Ctl.
    size_t checkpointBytes() const {
        return (size_t)N * (3 * sizeof(double2) + 1) + (size_t)E * 9 * sizeof(double) + sizeof(Ctl) + 64;
    }

    // One launch copies every array (k_ckpt_copy); the stream orders it against the steps.
    int checkpointCopy(Checkpoint& c, bool save) {
        CkptList list;
        size_t at = 0;
        int k = 0;
        auto region = [&](void* live, size_t bytes) {
            list.r[k++] = CkptRegion{(char*)live, at, bytes};
            at += (bytes + 15) & ~(size_t)15;
        };
        region(ctl, sizeof(Ctl));
        region(d.u, sizeof(double2) * N);
        region(d.v, sizeof(double2) * N);
        region(d.f, sizeof(double2) * N);
        region(d.sig0, sizeof(double) * E);
        region(d.sig1, sizeof(double) * E);
        region(d.sig2, sizeof(double) * E);
        region(d.sig3, sizeof(double) * E);
        region(d.epsP, sizeof(double) * E);
        region(d.work, sizeof(double) * E);
        region(d.hgR, sizeof(double) * E);
        region(d.hgZ, sizeof(double) * E);
        region(d.hgK, sizeof(double) * E);
        region(d.flags, N);
        const int grid = std::min(blocks_for((int)((N > E ? N : E) + 1)), 256);
        hipLaunchKernelGGL(k_ckpt_copy, dim3(grid), dim3(BLOCK), 0, stream, list, c.mem, save ? 1 : 0);
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    int saveCheckpoint(Checkpoint& c, long long steps) {
        if (!c.mem) HIP_TRY(hipMalloc((void**)&c.mem, checkpointBytes() + 16 * 16));
        c.steps = steps;
        c.q = q;
        return checkpointCopy(c, true);
    }

    int restoreCheckpoint(Checkpoint& c) {
        fastOpen = false;
        touch();
        TRY(checkpointCopy(c, false));
        q = c.q;
        return NF_OK;
    }

    // The forward run with checkpoints: the plain steps in chunks that end on multiples of
    // adjK, each chunk preceded by a checkpoint of the state it starts from.
    int adjointRun(double endTime, long long maxSteps, long long* taken) {
        Step h;
        TRY(readStep(h));
        const long long start = h.steps;
        if (taken) *taken = 0;
        const long long limitAbs = maxSteps > 0 ? start + maxSteps : 0x7FFFFFFFFFFFFFFFLL;
        if (h.time < endTime && maxSteps != 0) {
            TRY(adjointForward(endTime, start, limitAbs, h.steps == 0));
        }
        Step e;
        TRY(readStep(e));
        // Checkpoints enqueued past the run's end hold the final state under a step count it
        // never reached: drop them.
        while (!checkpoints.empty() && checkpoints.back().steps > e.steps) {
            if (checkpoints.back().mem) HIP_TRY(hipFree(checkpoints.back().mem));
            checkpoints.pop_back();
        }
        if (taken) *taken = e.steps - start;
        TRY(saveCheckpoint(finalState, e.steps));
        return sync();
    }

    // The chunks of the forward run, enqueued one ahead of the status the host reads: each
    // chunk is a checkpoint of the state it starts from (on a multiple of adjK), a launch that
    // sets the chunk's step limit on the device, the node launch that opens it (PHASE_RESUME:
    // a no-op once the run has ended), its steps, and the launches that close it.
    int adjointForward(double endTime, long long start, long long limitAbs, bool initial) {
        adjoint = false;  // the chunks run the plain (fused) steps
        struct Restore {
            nf_solver* s;
            ~Restore() { s->adjoint = true; s->fusedRun = false; }
        } restore{this};
        openElement = false;
        lastK3 = -1;
        fusedRun = fusedEligible();
        if (fusedRun) TRY(fusedBuffers());
        if (!statusHost) {
            HIP_TRY(hipHostMalloc((void**)&statusHost, 2 * sizeof(Step), hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&statusEvent[0], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&statusEvent[1], hipEventDisableTiming));
        }
        long long cur = start, pendingEnd = 0;
        int slot = 0;
        bool pending = false;
        while (true) {
            if (cur % adjK == 0 && (checkpoints.empty() || checkpoints.back().steps < cur)) {
                checkpoints.emplace_back();
                TRY(saveCheckpoint(checkpoints.back(), cur));
            }
            const long long chunkEnd = std::min<long long>(cur - cur % adjK + adjK, limitAbs);
            if (chunkEnd <= cur) break;
            hipLaunchKernelGGL(k_set_limits, dim3(1), dim3(1), 0, stream, ctl, q, endTime, chunkEnd);
            if (initial && cur == start) {
                TRY(launchElement(MODE_INITIAL, q));
                TRY(launchGN(PHASE_START, q, 1));
            } else {
                TRY(launchGN(PHASE_RESUME, q, 0));
            }
            q ^= 1;
            for (long long k = cur; k < chunkEnd; k++) TRY(enqueueStep());
            TRY(closeFused());
            HIP_TRY(hipMemcpyAsync(&statusHost[slot], &ctl->s[q], sizeof(Step), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(statusEvent[slot], stream));
            if (pending) {
                HIP_TRY(hipEventSynchronize(statusEvent[slot ^ 1]));
                const Step& st = statusHost[slot ^ 1];
                if (st.steps < pendingEnd || !(st.time < endTime) || st.steps >= limitAbs) break;
            }
            pending = true;
            pendingEnd = chunkEnd;
            slot ^= 1;
            cur = chunkEnd;
        }
        return sync();
    }

    int adjointAllocate() {
        if (adjAllocated) return NF_OK;
        ARec& a = ad.rec;
        const size_t K = adjK;
        a.K = adjK;
        TRY(alloc(&a.u, (K + 1) * N));
        TRY(alloc(&a.v, (K + 1) * N));
        TRY(alloc(&a.f, (K + 1) * N));
        TRY(alloc(&a.flags, (K + 1) * N));
        TRY(alloc(&a.sig, K * 4 * E));
        TRY(alloc(&a.ep, K * E));
        TRY(alloc(&a.work, K * E));
        TRY(alloc(&a.hg, K * 2 * E));
        TRY(alloc(&a.dGamma, K * E));
        TRY(alloc(&a.status, K * E));
        TRY(alloc(&a.capJac, K * CAP_DIRS * E));
        TRY(alloc(&a.edge, (size_t)E));
        TRY(alloc(&a.stepDt, K + 1));
        TRY(alloc(&a.stepDtPrev, K + 1));
        TRY(alloc(&a.minEdgeBits, K + 1));
        TRY(alloc(&a.argmin, K + 1));
        TRY(alloc(&a.capList, K * E));
        TRY(alloc(&a.capCount, 1));
        const size_t W = AW_MAX;
        TRY(alloc(&ad.uBar, W * N));
        TRY(alloc(&ad.vBar, W * N));
        TRY(alloc(&ad.fBar, W * N));
        TRY(alloc(&ad.cornerBar, W * 4 * E));
        TRY(alloc(&ad.sigBar, W * 4 * E));
        TRY(alloc(&ad.epBar, W * E));
        TRY(alloc(&ad.workBar, W * E));
        TRY(alloc(&ad.hgBar, W * 2 * E));
        // block counts for the widest layout (one objective per half workgroup)
        ad.elemBlocks = (E + BLOCK / 2 - 1) / (BLOCK / 2);
        ad.nodeBlocks = (N + BLOCK / 2 - 1) / (BLOCK / 2);
        TRY(alloc(&ad.elemPartial, W * ad.elemBlocks * A_EP));
        TRY(alloc(&ad.nodePartial, W * ad.nodeBlocks * A_NP));
        TRY(alloc(&ad.ctl, 1));
        adjAllocated = true;
        return NF_OK;
    }

    // Recomputes the steps after checkpoint c (count of them) with every step recorded; the
    // per-step dt, dtPrev and shortest edge come back in hDt, hDtPrev, hMin (index j).
    int adjointRecord(Checkpoint& c, int count, std::vector<double>& hDt, std::vector<double>& hDtPrev,
                      std::vector<double>& hMin) {
        TRY(restoreCheckpoint(c));
        ARec& a = ad.rec;
        Step h;
        TRY(readStep(h));
        h.endTime = HUGE_VAL;
        h.stepLimit = c.steps + count;
        TRY(writeStep(h));
        HIP_TRY(hipMemsetAsync(a.capCount, 0, sizeof(int), stream));
        const int gSave = blocks_for(std::max(N, E));
        hipLaunchKernelGGL(nfa::k_adj_save, dim3(gSave), dim3(BLOCK), 0, stream, d, a, 0, q);
        if (c.steps == 0) {
            TRY(launchElement(MODE_INITIAL, q));
            TRY(launchGN(PHASE_START, q, 1));
        } else {
            TRY(launchGN(PHASE_START, q, 0));
        }
        q ^= 1;
        // Two launches a step: k_element_rec also saves the step's record and takes the
        // previous step's argmin; the last step's argmin follows the loop.
        for (int j = 1; j <= count; j++) {
            hipLaunchKernelGGL(nfa::k_element_rec, dim3(gSave), dim3(BLOCK), 0, stream, d, a, j, q);
            TRY(launchGN(PHASE_STEP, q, 0));
            q ^= 1;
        }
        if (count > 0) hipLaunchKernelGGL(nfa::k_adj_argmin, dim3(blocks_for(E)), dim3(BLOCK), 0, stream, d, a, count, q ^ 1);
        HIP_TRY(hipGetLastError());
        hDt.assign(count + 1, 0.0);
        hDtPrev.assign(count + 1, 0.0);
        std::vector<unsigned long long> bits(count + 1);
        int caps = 0;
        HIP_TRY(hipMemcpyAsync(hDt.data(), a.stepDt, sizeof(double) * (count + 1), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(hDtPrev.data(), a.stepDtPrev, sizeof(double) * (count + 1), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(bits.data(), a.minEdgeBits, sizeof(unsigned long long) * (count + 1), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(&caps, a.capCount, sizeof(int), hipMemcpyDeviceToHost, stream));
        TRY(sync());
        hMin.assign(count + 1, 0.0);
        for (int j = 1; j <= count; j++) hMin[j] = nfm::from_bits(bits[j]);
        if (caps > 0) {
            const long long threads = (long long)caps * CAP_DIRS;
            hipLaunchKernelGGL(nfa::k_adj_cap, dim3((unsigned)((threads + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, stream,
                               d, a, caps);
            HIP_TRY(hipGetLastError());
        }
        adjCaps += caps;
        return NF_OK;
    }
    long long adjCaps = 0;  // element passes differentiated through the capped solve's Jacobian

    // The reverse sweep: seeds [W][4][N] are the objectives' adjoints of (ur, uz, vr, vz) at
    // the final state; grad [W][6] receives their gradients.
    int adjointSweep(int W, const double* seeds, double* grad) {
        TRY(adjointAllocate());
        const long long Nsteps = finalState.steps;
        ad.W = W;
        ad.elemBlocks = (E + BLOCK / W - 1) / (BLOCK / W);
        ad.nodeBlocks = (N + BLOCK / W - 1) / (BLOCK / W);
        // zero every adjoint, then the seeds
        HIP_TRY(hipMemsetAsync(ad.vBar, 0, sizeof(double2) * W * N, stream));
        HIP_TRY(hipMemsetAsync(ad.fBar, 0, sizeof(double2) * W * N, stream));
        HIP_TRY(hipMemsetAsync(ad.sigBar, 0, sizeof(double) * W * 4 * E, stream));
        HIP_TRY(hipMemsetAsync(ad.epBar, 0, sizeof(double) * W * E, stream));
        HIP_TRY(hipMemsetAsync(ad.workBar, 0, sizeof(double) * W * E, stream));
        HIP_TRY(hipMemsetAsync(ad.hgBar, 0, sizeof(double) * W * 2 * E, stream));
        HIP_TRY(hipMemsetAsync(ad.elemPartial, 0, sizeof(double) * W * ad.elemBlocks * A_EP, stream));
        HIP_TRY(hipMemsetAsync(ad.nodePartial, 0, sizeof(double) * W * ad.nodeBlocks * A_NP, stream));
        HIP_TRY(hipMemsetAsync(ad.ctl, 0, sizeof(AdjCtl), stream));
        {
            std::vector<double2> u((size_t)W * N), v((size_t)W * N);
            for (int w = 0; w < W; w++)
                for (int i = 0; i < N; i++) {
                    const double* sd = seeds + (size_t)w * 4 * N;
                    u[(size_t)w * N + i] = make_double2(sd[i], sd[N + i]);
                    v[(size_t)w * N + i] = make_double2(sd[2 * N + i], sd[3 * N + i]);
                }
            HIP_TRY(hipMemcpyAsync(ad.uBar, u.data(), sizeof(double2) * W * N, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipMemcpyAsync(ad.vBar, v.data(), sizeof(double2) * W * N, hipMemcpyHostToDevice, stream));
            TRY(sync());
        }
        adjCaps = 0;
        bool first = true;
        double nextDt = 0.0, nextDtPrev = 0.0;
        std::vector<double> hDt, hDtPrev, hMin;
        const int eb = ad.elemBlocks, nbk = ad.nodeBlocks;
        for (int c = (int)checkpoints.size() - 1; c >= 0; c--) {
            Checkpoint& ck = checkpoints[c];
            const long long end = c + 1 < (int)checkpoints.size() ? checkpoints[c + 1].steps : Nsteps;
            const int count = (int)(end - ck.steps);
            if (count <= 0) continue;
            TRY(adjointRecord(ck, count, hDt, hDtPrev, hMin));
            for (int j = count; j >= 1; j--) {
                hipLaunchKernelGGL(nfa::k_adj_reduce, dim3(1), dim3(BLOCK), 0, stream, d, ad, nextDt, nextDtPrev,
                                   first ? 0 : 1, hMin[j], 1);
                hipLaunchKernelGGL(nfa::k_adj_element, dim3(eb), dim3(BLOCK), 0, stream, d, ad, j);
                if (hasWall) hipLaunchKernelGGL(nfa::k_adj_node<true>, dim3(nbk), dim3(BLOCK), 0, stream, d, ad, j);
                else hipLaunchKernelGGL(nfa::k_adj_node<false>, dim3(nbk), dim3(BLOCK), 0, stream, d, ad, j);
                nextDt = hDt[j];
                nextDtPrev = hDtPrev[j];
                first = false;
            }
            HIP_TRY(hipGetLastError());
        }
        // fold the first step's partials into the gradient
        hipLaunchKernelGGL(nfa::k_adj_reduce, dim3(1), dim3(BLOCK), 0, stream, d, ad, nextDt, nextDtPrev,
                           first ? 0 : 1, 0.0, 0);
        HIP_TRY(hipGetLastError());
        AdjCtl h;
        HIP_TRY(hipMemcpyAsync(&h, ad.ctl, sizeof h, hipMemcpyDeviceToHost, stream));
        TRY(sync());
        for (int w = 0; w < W; w++)
            for (int k = 0; k < NP; k++) grad[w * NP + k] = h.grad[w][k];
        // leave the solver as the forward run left it
        TRY(restoreCheckpoint(finalState));
        return sync();
    }

    // run steps while time < endTime and fewer than maxSteps have been taken.
    int runSteps(double endTime, long long maxSteps, long long* taken) {
        if (adjoint) return adjointRun(endTime, maxSteps, taken);
        if (tangent && (hasContact || hasDamage || hasErosion))
            return fail(NF_E_STATE, "the sensitivity kernels do not carry contact, damage or erosion");
        if (taken) *taken = 0;
        touch();
        long long start = -1;  // known on the host, or -1: recorded by k_open_run
        const bool fast = fastOpen && !tangent;
        fastOpen = false;  // set again only by a call that completes
        openElement = false;
        openPending = false;
        lastK3 = -1;
        fusedRun = false;
        const bool fused = fusedEligible();
        if (fused) TRY(fusedBuffers());
        // Past step 0, a fused run's first launch opens the call's first step itself (k_step,
        // STEP_OPEN); otherwise a node launch opens it.
        bool opened = false;
        if (fast) {
            // Past step 0 with the control block as the last call left it: the opening is made
            // on the device. A call whose end has already passed runs as no-op launches.
            if (maxSteps == 0) return NF_OK;
            if (fused) {
                openPending = true;  // the opening k_step opens the call too (k_open_run's work)
            } else {
                hipLaunchKernelGGL(k_open_run, dim3(1), dim3(1), 0, stream, ctl, q, endTime, maxSteps);
                TRY(timed(0, [&](hipEvent_t a, hipEvent_t b) { return launchGN(PHASE_RESUME, q, 0, a, b); }));
                opened = true;
            }
        } else {
            Step h;
            {
                // The term-overflow flag reports on this call's steps alone.
                Ctl full;
                TRY(readCtl(full));
                if (full.cOverflow) {
                    full.cOverflow = 0;
                    TRY(writeCtl(full));
                }
                h = full.s[q];
            }
            start = h.steps;
            if (!(h.time < endTime) || maxSteps == 0) return NF_OK;
            h.endTime = endTime;
            h.stepLimit = maxSteps > 0 ? start + maxSteps : 0x7FFFFFFFFFFFFFFFLL;
            TRY(writeStep(h));
            // Open the first step of this call.
            opened = true;
            if (tangent) {
                TRY(launchTangent(h.steps == 0 ? MODE_INITIAL : -1, PHASE_START, h.steps == 0 ? 1 : 0));
            } else if (h.steps == 0) {
                TRY(launchElement(MODE_INITIAL, q));  // :945
                TRY(timed(0, [&](hipEvent_t a, hipEvent_t b) { return launchGN(PHASE_START, q, 1, a, b); }));
            } else if (fused) {
                openPending = true;  // the host's end test passed; the device's agrees
                opened = false;
            } else {
                TRY(timed(0, [&](hipEvent_t a, hipEvent_t b) { return launchGN(PHASE_START, q, 0, a, b); }));
            }
        }
        HIP_TRY(hipGetLastError());
        if (opened) q ^= 1;
        fusedRun = fused;
        openEnd = endTime;
        openMaxSteps = maxSteps;
        // A step count with no end time (or a short one) is enqueued whole. With an end time the
        // device's end test decides, and the host keeps only a block of steps ahead of it (below):
        // enqueuing a large step cap whole would leave every step past the end time as a no-op
        // launch.
        if (maxSteps > 0 && (!(endTime < HUGE_VAL) || maxSteps <= 2 * GRAPH_STEPS)) {
            // A known count: enqueue every step, synchronise once at the end.
            for (long long left = maxSteps; left > 0;) {
                const int n = (int)(left < GRAPH_STEPS ? left : GRAPH_STEPS);
                TRY(launchSteps(n));
                left -= n;
            }
        } else {
            // Up to an end time (and the step limit, if any): keep one block of steps in flight
            // beyond the one whose status is being read, so the GPU never idles while the host
            // decides. Blocks enqueued after the end test fails run as no-ops on the device.
            if (!statusHost) {
                HIP_TRY(hipHostMalloc((void**)&statusHost, 2 * sizeof(Step), hipHostMallocDefault));
                HIP_TRY(hipEventCreateWithFlags(&statusEvent[0], hipEventDisableTiming));
                HIP_TRY(hipEventCreateWithFlags(&statusEvent[1], hipEventDisableTiming));
            }
            int slot = 0;
            bool pending = false;
            while (true) {
                TRY(launchSteps(GRAPH_STEPS));
                HIP_TRY(hipMemcpyAsync(&statusHost[slot], &ctl->s[q], sizeof(Step), hipMemcpyDeviceToHost, stream));
                HIP_TRY(hipEventRecord(statusEvent[slot], stream));
                if (pending) {
                    HIP_TRY(hipEventSynchronize(statusEvent[slot ^ 1]));
                    if (!statusHost[slot ^ 1].go) break;
                }
                pending = true;
                slot ^= 1;
            }
        }
        TRY(closeFused());
        fusedRun = false;
        Ctl full;
        TRY(readCtl(full));
        const Step h = full.s[q];
        TRY(harvest());
        ctlCache = full;
        ctlCached = true;
        if (taken) *taken = h.steps - (start >= 0 ? start : full.runStart);
        fastOpen = !tangent && h.steps > 0;
        if (full.cOverflow)
            return fail(NF_E_STATE, "contact: a node received more than " + std::to_string(d.cMaxTerms) +
                                        " pair terms in one step");
        return NF_OK;
    }
};

#include "batch.h"
#include "defect_field.h"

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* nf_last_error(void) { return g_error.c_str(); }

int nf_material_check(const nf_material* m) {
    g_error.clear();
    return check_material(m);
}

int nf_solver_create(nf_solver** out, int32_t node_count, const double* r, const double* z,
                     int32_t element_count, const int32_t* conn, int32_t pressure_edge_count,
                     const int32_t* pressure_edges, int32_t material_count,
                     const nf_material* materials, const int32_t* material_of,
                     int32_t formulation, int32_t integration, int32_t kinematics,
                     double thickness, double cfl_safety) {
    g_error.clear();
    if (!out) return fail(NF_E_ARGUMENT, "out must not be null");
    *out = nullptr;
    if (node_count < 1 || element_count < 1 || !r || !z || !conn)
        return fail(NF_E_ARGUMENT, "mesh arrays must be non-empty");
    if (pressure_edge_count < 0 || (pressure_edge_count > 0 && !pressure_edges))
        return fail(NF_E_ARGUMENT, "bad pressure edges");
    if (material_count < 1 || !materials) return fail(NF_E_ARGUMENT, "need at least one material");
    if (!material_of && material_count != 1) return fail(NF_E_ARGUMENT, "a uniform mesh takes exactly one material");
    if (formulation < 0 || formulation > 2 || integration < 0 || integration > 1 || kinematics < 0 || kinematics > 1)
        return fail(NF_E_ARGUMENT, "unknown formulation, integration or kinematics");
    if ((long long)element_count * 4 + 2LL * pressure_edge_count > 0x7FFFFFFFLL)
        return fail(NF_E_ARGUMENT, "mesh too large for 32-bit slot indices");
    for (long long i = 0; i < 4LL * element_count; i++)
        if (conn[i] < 0 || conn[i] >= node_count) return fail(NF_E_ARGUMENT, "connectivity out of range");
    for (long long i = 0; i < 2LL * pressure_edge_count; i++)
        if (pressure_edges[i] < 0 || pressure_edges[i] >= node_count) return fail(NF_E_ARGUMENT, "pressure edge out of range");
    if (material_of)
        for (int e = 0; e < element_count; e++)
            if (material_of[e] < 0 || material_of[e] >= material_count) return fail(NF_E_ARGUMENT, "material index out of range");
    for (int m = 0; m < material_count; m++) {
        TRY(check_material(&materials[m]));
        if (formulation == NF_PLANE_STRESS && !std::isinf(materials[m].yield_stress))
            return fail(NF_E_ARGUMENT, "plane stress plasticity needs its own return map; not implemented");
    }
    // QuadMesh.checkOrientation mesh/QuadMesh.java:232-240
    for (int e = 0; e < element_count; e++) {
        double a = 0.0;
        for (int k = 0; k < 4; k++) {
            const int i = conn[e * 4 + k], j = conn[e * 4 + (k + 1) % 4];
            a += r[i] * z[j] - r[j] * z[i];
        }
        if (0.5 * a <= 0.0)
            return fail(NF_E_STATE, "element " + std::to_string(e) +
                                        " has non-positive area; node ordering must be counter-clockwise in (r, z)");
    }

    nf_solver* s = new nf_solver();
    auto bail = [&](int rc) {
        delete s;
        return rc;
    };
    int rc;
    if (const char* g = std::getenv("NEOFIZ_GRAPHS")) s->useGraphs = g[0] == '1';
    if (const char* g = std::getenv("NEOFIZ_FUSED")) s->fusedDisabled = g[0] == '0';
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NF_E_DEVICE, "hipStreamCreate failed: no usable GPU"));
    const int N = node_count, E = element_count;
    s->N = N;
    s->E = E;
    s->edges = pressure_edge_count;
    s->ppe = integration == NF_FULL ? 4 : 1;
    s->P = E * s->ppe;
    s->formulation = formulation;
    s->integration = integration;
    s->kinematics = kinematics;
    s->thickness = thickness;
    s->cfl = cfl_safety;
    s->r.assign(r, r + N);
    s->z.assign(z, z + N);
    s->conn.assign(conn, conn + 4 * (size_t)E);
    if (pressure_edge_count) s->pedges.assign(pressure_edges, pressure_edges + 2 * (size_t)pressure_edge_count);
    for (int m = 0; m < material_count; m++) s->materials.push_back({materials[m]});
    if (material_of) s->materialOf.assign(material_of, material_of + E);
    s->flagsHost.assign(N, 0);

    // Material table (solver/ExplicitSolver.java:289-299, J2.Flow.of).
    std::vector<Mat> table(material_count);
    for (int m = 0; m < material_count; m++) {
        table[m] = make_mat(s->materials[m].raw, formulation);
        if (m == 0) s->d.mat0 = table[0];
        if (table[m].hasJC) s->anyJC = true;
        if (table[m].hKin != 0.0) s->hasKin = true;
    }
    // Materials.fastestWave core/Materials.java:120-128
    if (s->materialOf.empty()) {
        s->fastestWave = s->materials[0].wave();
    } else {
        double fastest = 0.0;
        for (int e = 0; e < E; e++) {
            const double c = s->materials[s->materialOf[e]].wave();
            if (c > fastest) fastest = c;
        }
        s->fastestWave = fastest;
    }

    // CSR of force slots per node (buildNodeSlots :347-366): element corners in ascending
    // element order, then pressure edges. Slot index = 4 e + corner, or 4 (E + k) + end.
    s->slotStart.assign(N + 1, 0);
    for (size_t k = 0; k < s->conn.size(); k++) s->slotStart[s->conn[k] + 1]++;
    for (size_t k = 0; k < s->pedges.size(); k++) s->slotStart[s->pedges[k] + 1]++;
    for (int i = 0; i < N; i++) s->slotStart[i + 1] += s->slotStart[i];
    s->slotOf.assign(s->conn.size() + s->pedges.size(), 0);
    {
        std::vector<int> cursor(s->slotStart.begin(), s->slotStart.end() - 1);
        for (int e = 0; e < E; e++)
            for (int c = 0; c < 4; c++) s->slotOf[cursor[s->conn[e * 4 + c]]++] = 4 * e + c;
        for (int k = 0; k < pressure_edge_count; k++) {
            s->slotOf[cursor[s->pedges[2 * k]]++] = 4 * (E + k);
            s->slotOf[cursor[s->pedges[2 * k + 1]]++] = 4 * (E + k) + 1;
        }
    }

    Dev& d = s->d;
    d.N = N;
    d.E = E;
    d.P = s->P;
    d.edges = pressure_edge_count;
    d.ppe = s->ppe;
    d.axi = formulation == NF_AXISYMMETRIC;
    d.finite = kinematics == NF_FINITE_STRAIN;
    d.thickness = thickness;
    d.qhg = 0.05;  // DEFAULT_HOURGLASS_COEFFICIENT :59
    d.cfl = cfl_safety;
    d.fastestWave = s->fastestWave;
    d.massScale = 1.0;
    d.erosionDamage = HUGE_VAL;

    double2* x0 = nullptr;
    double* z0 = nullptr;
    double* mass = nullptr;
    int4* conn4 = nullptr;
    Mat* mats = nullptr;
    int* slotStart = nullptr;
    int4* slot4 = nullptr;
    int* slotOf = nullptr;
    if ((rc = s->alloc(&s->ctl, 1)) || (rc = s->alloc(&x0, N)) || (rc = s->alloc(&z0, N)) || (rc = s->alloc(&d.u, N)) ||
        (rc = s->alloc(&d.v, N)) || (rc = s->alloc(&d.f, N)) || (rc = s->alloc(&mass, N)) ||
        (rc = s->alloc(&d.flags, N)) || (rc = s->alloc(&conn4, E)) ||
        (rc = s->alloc(&mats, material_count)) || (rc = s->alloc(&d.sig0, s->P)) ||
        (rc = s->alloc(&d.sig1, s->P)) || (rc = s->alloc(&d.sig2, s->P)) || (rc = s->alloc(&d.sig3, s->P)) ||
        (rc = s->alloc(&d.back0, s->P)) || (rc = s->alloc(&d.back1, s->P)) ||
        (rc = s->alloc(&d.back2, s->P)) || (rc = s->alloc(&d.back3, s->P)) ||
        (rc = s->alloc(&d.epsP, s->P)) || (rc = s->alloc(&d.work, s->P)) ||
        (rc = s->alloc(&d.hgR, E)) || (rc = s->alloc(&d.hgZ, E)) || (rc = s->alloc(&d.hgK, E)) ||
        (rc = s->alloc(&d.slot, 4 * ((size_t)E + pressure_edge_count))) ||
        (rc = s->alloc(&slotStart, N + 1)) || (rc = s->alloc(&slotOf, s->slotOf.size())) ||
        (rc = s->alloc(&slot4, N)) ||
        (rc = s->alloc(&d.wallPartial, 3 * (size_t)blocks_for(N))) ||
        (rc = s->alloc(&d.wallCountPartial, 2 * (size_t)blocks_for(N))) ||
        (rc = s->alloc(&d.erodePartial, std::max((size_t)blocks_for(E + pressure_edge_count), ((size_t)E + 63) / 64))) ||
        (rc = s->alloc(&s->densityOfElement, E)))
        return bail(rc);
    d.ctl = s->ctl;
    d.x0 = x0;
    d.z0 = z0;
    d.mass = mass;
    d.conn = conn4;
    d.mats = mats;
    d.slotStart = slotStart;
    d.slotOf = slotOf;
    d.slot4 = slot4;
    if (pressure_edge_count) {
        int2* pe = nullptr;
        if ((rc = s->alloc(&pe, pressure_edge_count))) return bail(rc);
        d.pedges = pe;
        if (hipMemcpy(pe, s->pedges.data(), sizeof(int2) * pressure_edge_count, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(NF_E_DEVICE, "upload failed"));
    }
    if (material_of) {
        int* mi = nullptr;
        if ((rc = s->alloc(&mi, E))) return bail(rc);
        d.matIndex = mi;
        if (hipMemcpy(mi, material_of, sizeof(int) * E, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(NF_E_DEVICE, "upload failed"));
    }
    std::vector<int4> s4(N);
    for (int n = 0; n < N; n++) {
        int k[4] = {-1, -1, -1, -1};
        for (int c = 0; c < 4 && s->slotStart[n] + c < s->slotStart[n + 1]; c++) k[c] = s->slotOf[s->slotStart[n] + c];
        s4[n] = make_int4(k[0], k[1], k[2], k[3]);
    }
    std::vector<double2> x0h(N);
    for (int i = 0; i < N; i++) x0h[i] = make_double2(r[i], z[i]);
    std::vector<double> dens(E);
    for (int e = 0; e < E; e++) dens[e] = table[s->materialOf.empty() ? 0 : s->materialOf[e]].density;
    if (hipMemcpy(x0, x0h.data(), sizeof(double2) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(z0, z, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(conn4, conn, sizeof(int4) * E, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(mats, table.data(), sizeof(Mat) * material_count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(slotStart, s->slotStart.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(slot4, s4.data(), sizeof(int4) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(slotOf, s->slotOf.data(), sizeof(int) * s->slotOf.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->densityOfElement, dens.data(), sizeof(double) * E, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(NF_E_DEVICE, "upload failed"));
    hipLaunchKernelGGL(k_lump_mass, dim3(blocks_for(N)), dim3(BLOCK), 0, s->stream, d,
                       (const double*)s->densityOfElement, mass);
    if (hipGetLastError() != hipSuccess) return bail(fail(NF_E_DEVICE, "mass lumping launch failed"));

    // dt0 = cfl * minEdge / c (:334-338), QuadMesh.minimumEdgeLength :61-75
    double minEdge = 1.7976931348623157e308;
    for (int e = 0; e < E; e++)
        for (int k = 0; k < 4; k++) {
            const int i = conn[e * 4 + k], j = conn[e * 4 + (k + 1) % 4];
            const double dr = r[j] - r[i], dz = z[j] - z[i];
            const double len = std::sqrt(dr * dr + dz * dz);
            if (len < minEdge) minEdge = len;
        }
    Ctl h;
    std::memset(&h, 0, sizeof h);
    for (int k = 0; k < 2; k++) {
        h.s[k].dt = cfl_safety * minEdge / s->fastestWave;
        h.s[k].dtPrev = h.s[k].dt;
        h.s[k].minEdgeBits = MAX_BITS;
        h.fMin[k] = MAX_BITS;
    }
    s->dt0 = h.s[0].dt;
    if (E > 0) {
        // The element pass's patches (64 neighbours each, in the locality order) and its
        // first schedule, patch s in wave slot s.
        const std::vector<int> ord = element_order(E, conn, r, z, minEdge);
        const int P = (E + 63) / 64, nb = (P + 3) / 4;
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return bail(fail(NF_E_DEVICE, "device query failed"));
        const int units = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
        std::vector<int> padded((size_t)64 * P, -1), first((size_t)4 * nb, -1);
        std::copy(ord.begin(), ord.end(), padded.begin());
        for (int k = 0; k < P; k++) first[patch_slot(k, nb, units)] = k;
        std::vector<int4> corners(padded.size(), make_int4(0, 0, 0, 0));
        for (size_t k = 0; k < ord.size(); k++) {
            const int* c = conn + 4 * (size_t)ord[k];
            corners[k] = make_int4(c[0], c[1], c[2], c[3]);
        }
        int* o = nullptr;
        int4* oc = nullptr;
        int* po = nullptr;
        uint8_t* pc = nullptr;
        if ((rc = s->alloc(&o, padded.size())) || (rc = s->alloc(&oc, corners.size())) ||
            (rc = s->alloc(&po, 2 * first.size())) || (rc = s->alloc(&pc, 2 * (size_t)P)))
            return bail(rc);
        if (hipMemcpy(o, padded.data(), sizeof(int) * padded.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(oc, corners.data(), sizeof(int4) * corners.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(po, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(po + first.size(), first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(NF_E_DEVICE, "upload failed"));
        s->d.order = o;
        s->d.orderConn = oc;
        s->d.patchOf = po;
        s->d.patchCost = pc;
        s->d.patches = P;
        s->d.patchBlocks = nb;
        s->d.units = units;

        // The fused step kernel's node lists: every patch's corner nodes, each owned by the patch
        // that holds its lowest-numbered element. A node in no element or with more than four
        // force slots, or a patch with more than STEP_PATCH_NODES nodes, leaves the solver on the
        // two-launch step.
        std::vector<int> firstElem(N, -1);
        for (int e = E - 1; e >= 0; e--)
            for (int k = 0; k < 4; k++) firstElem[conn[4 * (size_t)e + k]] = e;
        std::vector<int> patchOfElem(E);
        for (size_t k = 0; k < ord.size(); k++) patchOfElem[ord[k]] = (int)(k / 64);
        bool fusable = N < (1 << 28);
        for (int i = 0; i < N && fusable; i++)
            fusable = firstElem[i] >= 0 && s->slotStart[i + 1] - s->slotStart[i] <= 4;
        // Each element slot's place in the node-major slots: 4 n + its rank in node n's gather.
        std::vector<int> nodeMajor(4 * (size_t)E, 0);
        for (int n = 0; n < N && fusable; n++)
            for (int j = s->slotStart[n]; j < s->slotStart[n + 1]; j++)
                if (s->slotOf[j] < 4 * E) nodeMajor[s->slotOf[j]] = 4 * n + (j - s->slotStart[n]);
        std::vector<int> pnStart(P + 1, 0), pnCode;
        std::vector<uchar4> local(padded.size(), make_uchar4(0, 0, 0, 0));
        std::vector<int4> dest(padded.size(), make_int4(0, 0, 0, 0));
        for (int p = 0; p < P && fusable; p++) {
            std::vector<int> nodes;
            for (int l = 0; l < 64; l++) {
                const int e = padded[64 * (size_t)p + l];
                if (e >= 0)
                    for (int k = 0; k < 4; k++) nodes.push_back(conn[4 * (size_t)e + k]);
            }
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            if ((int)nodes.size() > STEP_PATCH_NODES) {
                fusable = false;
                break;
            }
            for (int l = 0; l < 64; l++) {
                const int e = padded[64 * (size_t)p + l];
                if (e < 0) continue;
                unsigned char li[4];
                for (int k = 0; k < 4; k++)
                    li[k] = (unsigned char)(std::lower_bound(nodes.begin(), nodes.end(), conn[4 * (size_t)e + k]) - nodes.begin());
                local[64 * (size_t)p + l] = make_uchar4(li[0], li[1], li[2], li[3]);
                const size_t s0 = 4 * (size_t)e;
                dest[64 * (size_t)p + l] = make_int4(nodeMajor[s0], nodeMajor[s0 + 1], nodeMajor[s0 + 2], nodeMajor[s0 + 3]);
            }
            for (int node : nodes) {
                const int slots = s->slotStart[node + 1] - s->slotStart[node];
                pnCode.push_back(node | ((slots - 1) << 28) | (patchOfElem[firstElem[node]] == p ? (int)0x80000000u : 0));
            }
            pnStart[p + 1] = (int)pnCode.size();
        }
        if (fusable) {
            int* ps = nullptr;
            int* pc = nullptr;
            uchar4* ol = nullptr;
            int4* od = nullptr;
            if ((rc = s->alloc(&ps, pnStart.size())) || (rc = s->alloc(&pc, pnCode.size())) ||
                (rc = s->alloc(&ol, local.size())) || (rc = s->alloc(&od, dest.size())))
                return bail(rc);
            if (hipMemcpy(ps, pnStart.data(), sizeof(int) * pnStart.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(pc, pnCode.data(), sizeof(int) * pnCode.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(ol, local.data(), sizeof(uchar4) * local.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(od, dest.data(), sizeof(int4) * dest.size(), hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(NF_E_DEVICE, "upload failed"));
            // the first schedule's patch records, in both buffers
            std::vector<int4> runs(first.size(), make_int4(-1, 0, 0, 0));
            for (size_t k = 0; k < first.size(); k++)
                if (first[k] >= 0) runs[k] = make_int4(first[k], pnStart[first[k]], pnStart[first[k] + 1] - pnStart[first[k]], 0);
            int4* pr = nullptr;
            if ((rc = s->alloc(&pr, 2 * runs.size()))) return bail(rc);
            if (hipMemcpy(pr, runs.data(), sizeof(int4) * runs.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(pr + runs.size(), runs.data(), sizeof(int4) * runs.size(), hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(NF_E_DEVICE, "upload failed"));
            s->d.patchRun = pr;
            s->d.pnStart = ps;
            s->d.pnCode = pc;
            s->d.orderLocal = ol;
            s->d.orderDest = od;
        }
    }
    s->d.nodalBlocks = blocks_for(N);
    s->d.elementBlocks = blocks_for(E + pressure_edge_count);
    if ((rc = s->writeCtl(h))) return bail(rc);
    *out = s;
    return NF_OK;
}

int nf_solver_destroy(nf_solver* s) {
    delete s;
    return NF_OK;
}

int nf_set_threads(nf_solver* s, int32_t n) {
    s->touch();
    g_error.clear();
    if (n < 1) return fail(NF_E_ARGUMENT, "need at least one thread");
    return NF_OK;  // one GPU does the work; the answer never depended on the count
}

static int upload_flags(nf_solver* s) {
    // Keep the device's ON_WALL bits; replace the fixed bits.
    std::vector<uint8_t> cur(s->N);
    HIP_TRY(hipMemcpyAsync(cur.data(), s->d.flags, s->N, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->N; i++) cur[i] = (uint8_t)((cur[i] & ON_WALL) | s->flagsHost[i]);
    HIP_TRY(hipMemcpyAsync(s->d.flags, cur.data(), s->N, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NF_OK;
}

int nf_fix_radial(nf_solver* s, int32_t node) {
    s->touch();
    g_error.clear();
    if (node < 0 || node >= s->N) return fail(NF_E_ARGUMENT, "node out of range");
    s->flagsHost[node] |= FIXED_R;
    return upload_flags(s);
}
int nf_fix_axial(nf_solver* s, int32_t node) {
    s->touch();
    g_error.clear();
    if (node < 0 || node >= s->N) return fail(NF_E_ARGUMENT, "node out of range");
    s->flagsHost[node] |= FIXED_Z;
    return upload_flags(s);
}
int nf_fix_all_axial(nf_solver* s) {
    s->touch();
    for (auto& f : s->flagsHost) f |= FIXED_Z;
    return upload_flags(s);
}
int nf_set_pressure_ramp(nf_solver* s, double pascals, double ramp) {
    s->touch();
    s->d.targetPressure = pascals;
    s->d.rampDuration = ramp;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_relaxation_damping(nf_solver* s, double fraction, double omega) {
    s->touch();
    g_error.clear();
    if (omega <= 0.0) return fail(NF_E_ARGUMENT, "angular frequency must be positive");
    s->d.damping = 2.0 * fraction * omega;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_hourglass_coefficient(nf_solver* s, double q) {
    s->touch();
    g_error.clear();
    if (q < 0.0) return fail(NF_E_ARGUMENT, "hourglass coefficient must be >= 0");
    s->d.qhg = q;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_rigid_wall(nf_solver* s, int32_t enabled, double plane_z, double friction) {
    s->touch();
    g_error.clear();
    if (!enabled) {
        s->hasWall = false;
        s->d.wall = 0;
        s->graphDirty = true;
        return NF_OK;
    }
    if (friction < 0.0 || friction != friction) return fail(NF_E_ARGUMENT, "friction coefficient must be >= 0");
    s->hasWall = true;
    s->d.wall = 1;
    s->d.plane = plane_z;
    s->d.friction = friction;
    s->d.wallFrictional = friction > 0.0;
    s->graphDirty = true;
    // A new RigidWall starts with zero audits (solver/RigidWall.java:103-107).
    Ctl h;
    TRY(s->readCtl(h));
    h.wallImpulse = h.wallEnergyLoss = h.wallFriction = 0.0;
    h.wallContact = h.wallStuck = 0;
    return s->writeCtl(h);
}
// A solver carrying parameter sensitivities differentiates the Taylor step alone: contact,
// damage and erosion have no derivative kernels, so switching them on afterwards is refused
// (switching contact off is allowed: it is already off).
static int refuse_if_sensitive(const nf_solver* s, const char* what) {
    if (!s->tangent && !s->adjoint) return NF_OK;
    return fail(NF_E_STATE, std::string(what) + " cannot be set on a solver carrying sensitivities; "
                "they cover reduced integration of one Johnson-Cook material without damage, erosion or contact");
}

int nf_set_contact(nf_solver* s, int32_t enabled, double friction, double damping, double scale) {
    s->touch();
    g_error.clear();
    if (enabled) TRY(refuse_if_sensitive(s, "contact"));
    if (!enabled) {
        s->hasContact = false;
        s->d.contact = 0;
        s->graphDirty = true;
        return NF_OK;
    }
    if (friction < 0.0) return fail(NF_E_ARGUMENT, "friction must not be negative");
    if (damping < 0.0 || damping > 1.0) return fail(NF_E_ARGUMENT, "contact damping is a fraction of critical, so it lies in [0, 1]");
    if (!(scale > 0.0) || scale >= 2.0) return fail(NF_E_ARGUMENT, "the stiffness scale must lie in (0, 2)");
    Dev& d = s->d;
    const int N = s->N, E = s->E;
    if (!d.cDead) {
        // Static topology, built once: for every potential segment p = 4 e + k the other
        // potential segments with the same edge, and for every node the potential segments
        // touching it in ascending order.
        std::vector<std::pair<unsigned long long, int>> keyed(4 * (size_t)E);
        for (int p = 0; p < 4 * E; p++) {
            const int a = s->conn[p], b = s->conn[(p & ~3) + ((p + 1) & 3)];
            const unsigned lo = (unsigned)std::min(a, b), hi = (unsigned)std::max(a, b);
            keyed[p] = {((unsigned long long)lo << 32) | hi, p};
        }
        std::sort(keyed.begin(), keyed.end());
        std::vector<int> shareStart(4 * (size_t)E + 1, 0), shareOf;
        std::vector<std::vector<int>> shares(4 * (size_t)E);
        for (size_t i = 0; i < keyed.size();) {
            size_t j = i;
            while (j < keyed.size() && keyed[j].first == keyed[i].first) j++;
            for (size_t x = i; x < j; x++)
                for (size_t y = i; y < j; y++)
                    if (x != y) shares[keyed[x].second].push_back(keyed[y].second);
            i = j;
        }
        for (int p = 0; p < 4 * E; p++) {
            shareStart[p + 1] = shareStart[p] + (int)shares[p].size();
            for (int o : shares[p]) shareOf.push_back(o);
        }
        std::vector<int> nodeStart(N + 1, 0), nodeOf(8 * (size_t)E);
        for (int p = 0; p < 4 * E; p++) {
            nodeStart[s->conn[p] + 1]++;
            nodeStart[s->conn[(p & ~3) + ((p + 1) & 3)] + 1]++;
        }
        for (int n = 0; n < N; n++) nodeStart[n + 1] += nodeStart[n];
        {
            std::vector<int> cur(nodeStart.begin(), nodeStart.end() - 1);
            for (int p = 0; p < 4 * E; p++) {
                nodeOf[cur[s->conn[p]]++] = p;
                nodeOf[cur[s->conn[(p & ~3) + ((p + 1) & 3)]]++] = p;
            }
        }
        int* shareStartD = nullptr;
        int* shareOfD = nullptr;
        int* nodeStartD = nullptr;
        int* nodeOfD = nullptr;
        unsigned hmax = 64;
        while (hmax < 16u * (unsigned)E) hmax <<= 1;
        d.cApplyBlocks = std::min(blocks_for(N), CONTACT_GRID);
        TRY(s->alloc(&d.cDead, E));
        TRY(s->alloc(&d.cActive, 4 * (size_t)E));
        TRY(s->alloc(&d.cSurfFlag, N));
        TRY(s->alloc(&d.cPairOk, N));
        TRY(s->alloc(&d.cAct, 4 * (size_t)E));
        TRY(s->alloc(&d.cSurf, N));
        TRY(s->alloc(&shareStartD, 4 * (size_t)E + 1));
        TRY(s->alloc(&shareOfD, std::max<size_t>(1, shareOf.size())));
        TRY(s->alloc(&nodeStartD, N + 1));
        TRY(s->alloc(&nodeOfD, 8 * (size_t)E));
        TRY(s->alloc(&d.cLen, 4 * (size_t)E));
        TRY(s->alloc(&d.cNrm, 4 * (size_t)E));
        TRY(s->alloc(&d.cPn, N));
        TRY(s->alloc(&d.cLast, N));
        TRY(s->alloc(&d.cHead, hmax));
        TRY(s->alloc(&d.cNext, 8 * (size_t)E));
        TRY(s->alloc(&d.cEntCell, 8 * (size_t)E));
        TRY(s->alloc(&d.cPairHead, 4 * (size_t)E));
        TRY(s->alloc(&d.cPairNext, N));
        TRY(s->alloc(&d.cPair, (size_t)CPAIR * N));
        // Capture grid: entries for four cells per segment on average, at least 2^20.
        d.cFCap = (int)std::min<size_t>(0x7FFFFFFF, std::max<size_t>(1u << 20, 16 * (size_t)E));
        s->cFCapAlloc = d.cFCap;
        d.cMaxTerms = C_MAX_TERMS;
        // buckets for 4 per live segment (at most 4E segments), capped so that the heads
        // (C_SUB per bucket) stay within 1 GiB; past the cap buckets are shared more often
        d.cFHashMax = 64;
        while (d.cFHashMax < 16u * (unsigned)E && d.cFHashMax < (1u << 24)) d.cFHashMax <<= 1;
        TRY(s->alloc(&d.cFHead, (size_t)d.cFHashMax * C_SUB));
        TRY(s->alloc(&d.cFEnt, (size_t)d.cFCap));
        TRY(s->alloc(&d.cSegX, 4 * (size_t)E));
        // Capture cells twice the median surface segment of the reference mesh.
        {
            std::vector<double> lens;
            for (int p = 0; p < 4 * E; p++) {
                if (shareStart[p + 1] != shareStart[p]) continue;  // interior edge
                const int a = s->conn[p], b = s->conn[(p & ~3) + ((p + 1) & 3)];
                const double dr = s->r[b] - s->r[a], dz = s->z[b] - s->z[a];
                const double l = std::sqrt(dr * dr + dz * dz);
                if (l > 0.0) lens.push_back(l);
            }
            if (lens.empty()) {
                d.cFine = 1.0;
            } else {
                std::nth_element(lens.begin(), lens.begin() + lens.size() / 2, lens.end());
                d.cFine = 2.0 * lens[lens.size() / 2];
            }
        }
        TRY(s->alloc(&d.cPartial, 3 * (size_t)d.cApplyBlocks));
        TRY(s->alloc(&d.cPairsPartial, d.cApplyBlocks));
        HIP_TRY(hipMemcpy(shareStartD, shareStart.data(), sizeof(int) * shareStart.size(), hipMemcpyHostToDevice));
        if (!shareOf.empty())
            HIP_TRY(hipMemcpy(shareOfD, shareOf.data(), sizeof(int) * shareOf.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(nodeStartD, nodeStart.data(), sizeof(int) * nodeStart.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(nodeOfD, nodeOf.data(), sizeof(int) * nodeOf.size(), hipMemcpyHostToDevice));
        d.cShareStart = shareStartD;
        d.cShareOf = shareOfD;
        d.cNodeSegStart = nodeStartD;
        d.cNodeSegOf = nodeOfD;
        // Fixed-width copies of the two CSRs for the common case, so that a test costs one
        // dependent load less.
        std::vector<int> partner(4 * (size_t)E);
        for (int p = 0; p < 4 * E; p++) {
            const int c = shareStart[p + 1] - shareStart[p];
            partner[p] = c == 0 ? -1 : (c == 1 ? shareOf[shareStart[p]] : -2);
        }
        std::vector<int> seg8(8 * (size_t)N, -1);
        for (int n = 0; n < N; n++)
            for (int k = nodeStart[n]; k < nodeStart[n + 1] && k < nodeStart[n] + 8; k++)
                seg8[8 * (size_t)n + (k - nodeStart[n])] = nodeOf[k];
        int* partnerD = nullptr;
        int4* seg8D = nullptr;
        TRY(s->alloc(&partnerD, 4 * (size_t)E));
        TRY(s->alloc(&seg8D, 2 * (size_t)N));
        HIP_TRY(hipMemcpy(partnerD, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(seg8D, seg8.data(), sizeof(int) * seg8.size(), hipMemcpyHostToDevice));
        d.cPartner = partnerD;
        d.cNodeSeg8 = seg8D;
    }
    // new Contact(mesh) :273-280: nothing forgotten, no history, audits from zero.
    HIP_TRY(hipMemsetAsync(d.cDead, 0, E, s->stream));
    Ctl h;
    TRY(s->readCtl(h));
    h.surfaceStale = 1;
    h.cNAct = 0;
    h.cNSurf = 0;
    h.cPairs = 0;
    h.cOverflow = 0;
    h.cEnergy = 0.0;
    h.cDissipation = 0.0;
    h.cMaxPen = 0.0;
    h.cEscaped = 0;
    TRY(s->writeCtl(h));
    d.contact = 1;
    d.cFriction = friction;
    d.cDamping = damping;
    d.cScale = scale;
    d.cRef = s->dt0;
    hipLaunchKernelGGL(k_c_rebuild, dim3(blocks_for_chunks(4 * E, C_CHUNK_SEG)), dim3(BLOCK), 0, s->stream, d, s->q, 0);
    const int nb = blocks_for_chunks(N, C_CHUNK_NODE);  // the node part alone: no geometry yet
    hipLaunchKernelGGL(k_c_nodes_geometry, dim3(nb), dim3(BLOCK), 0, s->stream, d, s->q, 0, nb);
    HIP_TRY(hipGetLastError());
    // The lists are built: clear the flag so the first assembly does not append them again.
    TRY(s->sync());
    TRY(s->readCtl(h));
    h.surfaceStale = 0;
    TRY(s->writeCtl(h));
    s->hasContact = true;
    s->graphDirty = true;
    return NF_OK;
}
int nf_contact_set_limits(nf_solver* s, int32_t capture_capacity, int32_t max_terms) {
    s->touch();
    g_error.clear();
    if (!s->hasContact) return fail(NF_E_STATE, "contact is not set");
    if (max_terms == 0 || max_terms > C_MAX_TERMS || max_terms < -1)
        return fail(NF_E_ARGUMENT, "max_terms lies in [1, " + std::to_string(C_MAX_TERMS) + "], or -1 for the default");
    if (capture_capacity < -1) return fail(NF_E_ARGUMENT, "capture_capacity is a count, or -1 for the default");
    s->d.cFCap = capture_capacity < 0 ? s->cFCapAlloc : std::min(capture_capacity, s->cFCapAlloc);
    s->d.cMaxTerms = max_terms < 0 ? C_MAX_TERMS : max_terms;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_gravity(nf_solver* s, double ar, double az) {
    s->touch();
    s->d.gravityR = ar;
    s->d.gravityZ = az;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_damage(nf_solver* s, const double* gf, const double* eps) {  // :729-791
    g_error.clear();
    TRY(refuse_if_sensitive(s, "damage"));
    if (!gf || !eps) return fail(NF_E_ARGUMENT, "fracture energy and failure strain are required");
    for (const HostMaterial& m : s->materials) {
    s->touch();
        const double worst = 2.0 * m.mu() + (2.0 / 3.0) * (m.raw.kinematic_hardening - m.raw.youngs_modulus);
        if (!(worst > 0.0)) return fail(NF_E_ARGUMENT, "a fully clamped softening modulus would make the return map singular");
    }
    const int E = s->E, ppe = s->ppe, P = s->P;
    std::vector<double> fe(P), ym(P), fs(P), geom(5 * (size_t)P);
    for (int e = 0; e < E; e++) {
        const int* c = &s->conn[4 * e];
        const double uR = 0.5 * (s->r[c[1]] - s->r[c[0]] + s->r[c[2]] - s->r[c[3]]);
        const double uZ = 0.5 * (s->z[c[1]] - s->z[c[0]] + s->z[c[2]] - s->z[c[3]]);
        const double vR = 0.5 * (s->r[c[3]] - s->r[c[0]] + s->r[c[2]] - s->r[c[1]]);
        const double vZ = 0.5 * (s->z[c[3]] - s->z[c[0]] + s->z[c[2]] - s->z[c[1]]);
        double a = 0.0;
        for (int k = 0; k < 4; k++) {
            const int i = c[k], j = c[(k + 1) % 4];
            a += s->r[i] * s->z[j] - s->r[j] * s->z[i];
        }
        const double area = std::fabs(0.5 * a);
        for (int g = 0; g < ppe; g++) {
            const int p = e * ppe + g;
            fs[p] = eps[e];
            fe[p] = gf[e];
            ym[p] = s->matE(e);
            geom[5 * (size_t)p] = area;
            geom[5 * (size_t)p + 1] = uR;
            geom[5 * (size_t)p + 2] = uZ;
            geom[5 * (size_t)p + 3] = vR;
            geom[5 * (size_t)p + 4] = vZ;
        }
    }
    for (int p = 0; p < P; p++) {
        if (!(fe[p] > 0.0)) return fail(NF_E_ARGUMENT, "fracture energy must be positive");
        if (!(ym[p] > 0.0)) return fail(NF_E_ARGUMENT, "Young's modulus must be positive");
    }
    for (int p = 0; p < P; p++) {
        if (!(fs[p] > 0.0)) return fail(NF_E_ARGUMENT, "failure strain must be positive");
        if (!(geom[5 * (size_t)p] > 0.0)) return fail(NF_E_ARGUMENT, "area must be positive");
    }
    Dev& d = s->d;
    if (!d.dStrength) {
        double *a, *b, *c2, *f1, *f2, *f3, *g;
        TRY(s->alloc(&a, P));
        TRY(s->alloc(&b, P));
        TRY(s->alloc(&c2, P));
        TRY(s->alloc(&f1, P));
        TRY(s->alloc(&f2, P));
        TRY(s->alloc(&f3, P));
        TRY(s->alloc(&g, 5 * (size_t)P));
        d.dStrength = a; d.dSoftening = b; d.dWidth = c2;
        d.dFailure = f1; d.dGf = f2; d.dYoungs = f3; d.dGeom = g;
    } else {
        HIP_TRY(hipMemsetAsync(d.dStrength, 0, sizeof(double) * P, s->stream));
        HIP_TRY(hipMemsetAsync(d.dSoftening, 0, sizeof(double) * P, s->stream));
        HIP_TRY(hipMemsetAsync(d.dWidth, 0, sizeof(double) * P, s->stream));
    }
    HIP_TRY(hipMemcpyAsync((void*)d.dFailure, fs.data(), sizeof(double) * P, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync((void*)d.dGf, fe.data(), sizeof(double) * P, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync((void*)d.dYoungs, ym.data(), sizeof(double) * P, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync((void*)d.dGeom, geom.data(), sizeof(double) * 5 * P, hipMemcpyHostToDevice, s->stream));
    TRY(s->sync());
    s->hasDamage = true;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_erosion(nf_solver* s, double threshold) {  // :549-561
    s->touch();
    g_error.clear();
    TRY(refuse_if_sensitive(s, "erosion"));
    if (!s->hasDamage) return fail(NF_E_STATE, "erosion needs a failure model; call setDamage first");
    if (!(threshold > 0.0) || threshold > 1.0)
        return fail(NF_E_ARGUMENT, "the erosion threshold is a damage, so it lies in (0, 1]");
    s->d.erosionDamage = threshold;
    if (!s->d.gone) {
    s->touch();
        uint8_t* g = nullptr;
        TRY(s->alloc(&g, s->E));
        s->d.gone = g;
    }
    s->hasErosion = true;
    s->graphDirty = true;
    return NF_OK;
}
int nf_set_mass_scaling(nf_solver* s, double target) {  // :601-619
    s->touch();
    g_error.clear();
    if (!(target > 0.0)) return fail(NF_E_ARGUMENT, "the target step must be positive");
    Step h;
    TRY(s->readStep(h));
    if (target < h.dt) return fail(NF_E_ARGUMENT, "mass scaling adds mass and cannot shorten a step");
    const double factor = (target / h.dt) * (target / h.dt);
    hipLaunchKernelGGL(k_scale_mass, dim3(blocks_for(s->N)), dim3(BLOCK), 0, s->stream,
                       (double*)s->d.mass, s->N, factor);
    HIP_TRY(hipGetLastError());
    s->massScale *= factor;
    s->d.massScale = s->massScale;
    h.dt = target;
    h.dtPrev = target;
    s->dt0 = target;
    s->d.cRef = target;
    s->graphDirty = true;
    return s->writeStep(h);
}
int nf_set_velocity(nf_solver* s, int32_t node, double vr, double vz) {
    s->touch();
    g_error.clear();
    if (node < 0 || node >= s->N) return fail(NF_E_ARGUMENT, "node out of range");
    const double2 v = make_double2(vr, vz);
    HIP_TRY(hipMemcpyAsync(s->d.v + node, &v, sizeof v, hipMemcpyHostToDevice, s->stream));
    return s->sync();
}
int nf_set_uniform_velocity(nf_solver* s, double vr, double vz) {
    s->touch();
    hipLaunchKernelGGL(k_fill2, dim3(blocks_for(s->N)), dim3(BLOCK), 0, s->stream, s->d.v, s->N, vr, vz);
    HIP_TRY(hipGetLastError());
    return s->sync();
}
int nf_run(nf_solver* s, int64_t steps) {
    s->touch();
    g_error.clear();
    if (steps <= 0) return NF_OK;
    long long taken = 0;
    return s->runSteps(HUGE_VAL, steps, &taken);
}
int nf_run_until(nf_solver* s, double end_time, int64_t max_steps, int64_t* steps_taken) {
    s->touch();
    g_error.clear();
    long long taken = 0;
    TRY(s->runSteps(end_time, max_steps > 0 ? max_steps : -1, &taken));
    if (steps_taken) *steps_taken = taken;
    return NF_OK;
}
int nf_advance(nf_solver* s) {
    s->touch();
    g_error.clear();
    hipLaunchKernelGGL(k_advance_disp, dim3(blocks_for(s->N)), dim3(BLOCK), 0, s->stream, s->d, s->q);
    HIP_TRY(hipGetLastError());
    TRY(s->assembleAdvance());
    return s->sync();
}

int nf_get_node_field(nf_solver* s, int32_t field, double* out) {
    g_error.clear();
    const int N = s->N;
    if (field < 0 || field >= NF_NODE_FIELD_COUNT) return fail(NF_E_ARGUMENT, "unknown node field");
    if (field == NF_NODE_R || field == NF_NODE_Z) {
        std::memcpy(out, field == NF_NODE_R ? s->r.data() : s->z.data(), sizeof(double) * N);
        return NF_OK;
    }
    if (field == NF_NODE_MASS) {
        HIP_TRY(hipMemcpyAsync(out, s->d.mass, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
        return s->sync();
    }
    if (field >= NF_NODE_ON_WALL) {
        std::vector<uint8_t> fl(N);
        HIP_TRY(hipMemcpyAsync(fl.data(), s->d.flags, N, hipMemcpyDeviceToHost, s->stream));
        TRY(s->sync());
        const uint8_t bit = field == NF_NODE_ON_WALL ? ON_WALL : field == NF_NODE_FIXED_R ? FIXED_R : FIXED_Z;
        for (int i = 0; i < N; i++) out[i] = (fl[i] & bit) ? 1.0 : 0.0;
        return NF_OK;
    }
    const double2* src = field <= NF_NODE_UZ ? s->d.u : field <= NF_NODE_VZ ? s->d.v : s->d.f;
    std::vector<double2> tmp(N);
    HIP_TRY(hipMemcpyAsync(tmp.data(), src, sizeof(double2) * N, hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    const bool second = (field & 1) != 0;  // UR/VR/FR even, UZ/VZ/FZ odd
    for (int i = 0; i < N; i++) out[i] = second ? tmp[i].y : tmp[i].x;
    return NF_OK;
}

int nf_put_node_field(nf_solver* s, int32_t field, const double* in) {
    s->touch();
    g_error.clear();
    if (field < NF_NODE_UR || field > NF_NODE_VZ) return fail(NF_E_ARGUMENT, "only UR, UZ, VR and VZ may be written");
    double2* dst = field <= NF_NODE_UZ ? s->d.u : s->d.v;
    std::vector<double2> tmp(s->N);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dst, sizeof(double2) * s->N, hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    const bool second = (field & 1) != 0;
    for (int i = 0; i < s->N; i++) (second ? tmp[i].y : tmp[i].x) = in[i];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), sizeof(double2) * s->N, hipMemcpyHostToDevice, s->stream));
    return s->sync();
}

int nf_get_point_field(nf_solver* s, int32_t field, double* out) {
    g_error.clear();
    const int P = s->P;
    const Dev& d = s->d;
    auto copy = [&](const double* src, double* dst) -> int {
        HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * P, hipMemcpyDeviceToHost, s->stream));
        return NF_OK;
    };
    if (field == NF_POINT_STRESS || field == NF_POINT_BACK) {
        const bool st = field == NF_POINT_STRESS;
        std::vector<double> c0(P), c1(P), c2(P), c3(P);
        TRY(copy(st ? d.sig0 : d.back0, c0.data()));
        TRY(copy(st ? d.sig1 : d.back1, c1.data()));
        TRY(copy(st ? d.sig2 : d.back2, c2.data()));
        TRY(copy(st ? d.sig3 : d.back3, c3.data()));
        TRY(s->sync());
        for (int p = 0; p < P; p++) {
            out[4 * p] = c0[p];
            out[4 * p + 1] = c1[p];
            out[4 * p + 2] = c2[p];
            out[4 * p + 3] = c3[p];
        }
        return NF_OK;
    }
    if (field == NF_POINT_TEMPERATURE) {  // JohnsonCook.temperature, core/JohnsonCook.java:218-221
        std::vector<double> work(P);
        TRY(copy(d.work, work.data()));
        TRY(s->sync());
        for (int p = 0; p < P; p++) {
            const HostMaterial& m = s->materials[s->materialOf.empty() ? 0 : s->materialOf[p / s->ppe]];
            const nf_johnson_cook& j = m.raw.johnson_cook;
            out[p] = m.raw.has_johnson_cook ? j.room_temperature + j.taylor_quinney * work[p] / (m.raw.density * j.specific_heat)
                                            : NAN;
        }
        return NF_OK;
    }
    if (field == NF_POINT_DAMAGE) {  // Damage.at, solver/Damage.java:342-347
        if (!d.dStrength) {
            std::memset(out, 0, sizeof(double) * P);
            return NF_OK;
        }
        std::vector<double> st(P), so(P), fs(P), ep(P);
        TRY(copy(d.dStrength, st.data()));
        TRY(copy(d.dSoftening, so.data()));
        TRY(copy(d.dFailure, fs.data()));
        TRY(copy(d.epsP, ep.data()));
        TRY(s->sync());
        for (int p = 0; p < P; p++) {
            double v = 0.0;
            if (st[p] > 0.0) {
                const double dd = -so[p] * (ep[p] - fs[p]) / st[p];
                v = dd <= 0.0 ? 0.0 : (dd >= 1.0 ? 1.0 : dd);
            }
            out[p] = v;
        }
        return NF_OK;
    }
    const double* src = nullptr;
    switch (field) {
        case NF_POINT_PLASTIC_STRAIN: src = d.epsP; break;
        case NF_POINT_PLASTIC_WORK: src = d.work; break;
        case NF_POINT_DAMAGE_STRENGTH: src = d.dStrength; break;
        case NF_POINT_DAMAGE_SOFTENING: src = d.dSoftening; break;
        case NF_POINT_DAMAGE_WIDTH: src = d.dWidth; break;
        case NF_POINT_DAMAGE_FAILURE_STRAIN: src = d.dFailure; break;
        case NF_POINT_DAMAGE_FRACTURE_ENERGY: src = d.dGf; break;
        default: return fail(NF_E_ARGUMENT, "unknown point field");
    }
    if (!src) {
        std::memset(out, 0, sizeof(double) * P);
        return NF_OK;
    }
    TRY(copy(src, out));
    return s->sync();
}

int nf_get_element_field(nf_solver* s, int32_t field, double* out) {
    g_error.clear();
    const int E = s->E;
    const double* src = nullptr;
    switch (field) {
        case NF_ELEMENT_HG_R: src = s->d.hgR; break;
        case NF_ELEMENT_HG_Z: src = s->d.hgZ; break;
        case NF_ELEMENT_HG_STIFFNESS: src = s->d.hgK; break;
        case NF_ELEMENT_ERODED: {
            if (!s->d.gone) {
                std::memset(out, 0, sizeof(double) * E);
                return NF_OK;
            }
            std::vector<uint8_t> g(E);
            HIP_TRY(hipMemcpyAsync(g.data(), s->d.gone, E, hipMemcpyDeviceToHost, s->stream));
            TRY(s->sync());
            for (int e = 0; e < E; e++) out[e] = g[e] ? 1.0 : 0.0;
            return NF_OK;
        }
        default: return fail(NF_E_ARGUMENT, "unknown element field");
    }
    HIP_TRY(hipMemcpyAsync(out, src, sizeof(double) * E, hipMemcpyDeviceToHost, s->stream));
    return s->sync();
}

int nf_film_capture(nf_solver* s, int32_t field_count, const int32_t* fields, double* r, double* z,
                    double* values) {
    g_error.clear();
    if (!s) return fail(NF_E_ARGUMENT, "null handle");
    s->touch();
    if (field_count < 0 || field_count > FILM_MAX)
        return fail(NF_E_ARGUMENT, "a frame carries 0 to " + std::to_string(FILM_MAX) + " fields");
    if (!r || !z || (field_count && (!fields || !values))) return fail(NF_E_ARGUMENT, "null array");
    FilmFields ff{};
    ff.n = field_count;
    for (int f = 0; f < field_count; f++) {
        if (fields[f] < NF_FILM_PLASTIC_STRAIN || fields[f] > NF_FILM_ALIVE)
            return fail(NF_E_ARGUMENT, "unknown film field " + std::to_string(fields[f]));
        ff.id[f] = fields[f];
    }
    const size_t N = (size_t)s->N, E = (size_t)s->E;
    const size_t need = 2 * N + (size_t)field_count * E;
    if (need > s->filmCap) {
        if (s->filmBuf) HIP_TRY(hipFree(s->filmBuf));
        s->filmBuf = nullptr;
        s->filmCap = 0;
        HIP_TRY(hipMalloc((void**)&s->filmBuf, sizeof(double) * need));
        s->filmCap = need;
    }
    const int n = (int)std::max(N, E);
    hipLaunchKernelGGL(k_film, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->d, ff, s->filmBuf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(r, s->filmBuf, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(z, s->filmBuf + N, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
    if (field_count)
        HIP_TRY(hipMemcpyAsync(values, s->filmBuf + 2 * N, sizeof(double) * field_count * E,
                               hipMemcpyDeviceToHost, s->stream));
    return s->sync();
}

int nf_get_state(nf_solver* s, nf_state* o) {
    g_error.clear();
    Ctl h;
    TRY(s->readCtl(h));
    const Step& st = h.s[s->q];
    std::memset(o, 0, sizeof *o);
    o->time = st.time;
    o->dt = st.dt;
    o->dt_previous = st.dtPrev;
    o->dt0 = s->dt0;
    o->mass_scale = s->massScale;
    o->steps = st.steps;
    o->node_count = s->N;
    o->element_count = s->E;
    o->points_per_element = s->ppe;
    o->eroded_elements = h.erodedCount;
    o->inverted_elements = h.invertedCount;
    o->eroded_energy = h.erodedEnergy;
    if (s->hasWall) {
        o->wall_contact_count = h.wallContact;
        o->wall_stuck_count = h.wallStuck;
        o->wall_impulse = h.wallImpulse;
        o->wall_energy_loss = h.wallEnergyLoss;
        o->wall_friction_dissipation = h.wallFriction;
    }
    if (s->hasContact) {
        o->contact_pairs = h.cPairs;
        o->contact_energy = h.cEnergy;
        o->contact_dissipation = h.cDissipation;
        o->contact_max_penetration = h.cMaxPen;
        o->contact_escaped_nodes = h.cEscaped;
    }
    return NF_OK;
}

// The contact's pairs of the last step, in surface-node order, as (node, a, b).
int nf_contact_pairs(nf_solver* s, int32_t capacity, int32_t* out, int32_t* count) {
    s->touch();
    g_error.clear();
    if (!count) return fail(NF_E_ARGUMENT, "count must not be null");
    *count = 0;
    if (!s->hasContact) return NF_OK;
    std::vector<int> last(s->N);
    std::vector<uint8_t> on(s->N);
    HIP_TRY(hipMemcpyAsync(last.data(), s->d.cLast, sizeof(int) * s->N, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(on.data(), s->d.cSurfFlag, s->N, hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    int k = 0;
    for (int n = 0; n < s->N; n++) {
        if (!on[n] || last[n] < 0) continue;
        const int p = last[n];
        if (k < capacity && out) {
            out[3 * k] = n;
            out[3 * k + 1] = s->conn[p];
            out[3 * k + 2] = s->conn[(p & ~3) + ((p + 1) & 3)];
        }
        k++;
    }
    *count = k;
    return NF_OK;
}

// The contact's current surface segments (a, b) in surface order.
int nf_contact_surface(nf_solver* s, int32_t capacity, int32_t* out, int32_t* count) {
    s->touch();
    g_error.clear();
    if (!count) return fail(NF_E_ARGUMENT, "count must not be null");
    *count = 0;
    if (!s->hasContact) return NF_OK;
    std::vector<uint8_t> live(4 * (size_t)s->E);
    HIP_TRY(hipMemcpyAsync(live.data(), s->d.cActive, live.size(), hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    int k = 0;
    for (int p = 0; p < 4 * s->E; p++) {
        if (!live[p]) continue;
        if (k < capacity && out) {
            out[2 * k] = s->conn[p];
            out[2 * k + 1] = s->conn[(p & ~3) + ((p + 1) & 3)];
        }
        k++;
    }
    *count = k;
    return NF_OK;
}

// ExplicitSolver's single-value queries (solver/ExplicitSolver.java:826, :1778-1794,
// :1832-1840, :1867-1877, :1928-1930), from the device state copied out.
int nf_temperature(nf_solver* s, int32_t point, double* out) {
    s->touch();
    g_error.clear();
    if (point < 0 || point >= s->P) return fail(NF_E_ARGUMENT, "point out of range");
    const HostMaterial& m = s->materials[s->materialOf.empty() ? 0 : s->materialOf[point / s->ppe]];
    if (!m.raw.has_johnson_cook) return fail(NF_E_STATE, "this material has no thermal model; temperature is not defined");
    double w = 0.0;
    HIP_TRY(hipMemcpyAsync(&w, s->d.work + point, sizeof w, hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    const nf_johnson_cook& j = m.raw.johnson_cook;
    *out = j.room_temperature + j.taylor_quinney * w / (m.raw.density * j.specific_heat);
    return NF_OK;
}

int nf_plastic_front_radius(nf_solver* s, double* out) {
    s->touch();
    g_error.clear();
    std::vector<double> ep(s->P);
    TRY(nf_get_point_field(s, NF_POINT_PLASTIC_STRAIN, ep.data()));
    double front = 0.0;
    for (int e = 0; e < s->E; e++) {
        bool yielded = false;
        for (int g = 0; g < s->ppe; g++)
            if (ep[e * s->ppe + g] > 0.0) yielded = true;
        if (!yielded) continue;
        const int* c = &s->conn[4 * e];  // outerRadius :1932-1936, reference coordinates
        const double outer = std::max(std::max(s->r[c[0]], s->r[c[1]]), std::max(s->r[c[2]], s->r[c[3]]));
        front = std::max(front, outer);
    }
    *out = front;
    return NF_OK;
}

int nf_most_damaged_element(nf_solver* s, int32_t* out) {
    s->touch();
    g_error.clear();
    *out = -1;
    if (!s->hasDamage) return NF_OK;
    std::vector<double> dmg(s->P);
    TRY(nf_get_point_field(s, NF_POINT_DAMAGE, dmg.data()));
    double max = 0.0;
    for (int e = 0; e < s->E; e++)
        for (int g = 0; g < s->ppe; g++) {
            const double v = dmg[e * s->ppe + g];
            if (v > max) {
                max = v;
                *out = e;
            }
        }
    return NF_OK;
}

int nf_bore_pressure(nf_solver* s, double* out) {  // pressureAt(time) :1203-1207
    g_error.clear();
    Step h;
    TRY(s->readStep(h));
    const Dev& d = s->d;
    if (d.rampDuration <= 0.0) {
    s->touch();
        *out = d.targetPressure;
        return NF_OK;
    }
    const double t = h.time / d.rampDuration;
    const double sm = t != t ? t : (t < 1.0 ? t : 1.0);  // Math.min(1.0, t)
    *out = d.targetPressure * sm * sm * (3.0 - 2.0 * sm);
    return NF_OK;
}

int nf_mesh_z(nf_solver* s, int32_t node, double* out) {
    s->touch();
    g_error.clear();
    if (node < 0 || node >= s->N) return fail(NF_E_ARGUMENT, "node out of range");
    *out = s->z[node];
    return NF_OK;
}

int nf_set_damage_uniform(nf_solver* s, double fracture_energy, const double* failure_strain) {  // :713-717
    g_error.clear();
    const std::vector<double> gf(s->E, fracture_energy);
    return nf_set_damage(s, gf.data(), failure_strain);
}

// Audits from the device state, in the reference's summation order
// (solver/ExplicitSolver.java:875-886, :1610-1860).
int nf_get_audits(nf_solver* s, nf_audits* o) {
    s->touch();
    g_error.clear();
    std::memset(o, 0, sizeof *o);
    const int N = s->N, E = s->E, ppe = s->ppe, P = s->P;
    std::vector<double> mass(N), ur(N), uz(N), vr(N), vz(N), fr(N), fz(N), onw(N), fxr(N), fxz(N);
    TRY(nf_get_node_field(s, NF_NODE_MASS, mass.data()));
    TRY(nf_get_node_field(s, NF_NODE_VR, vr.data()));
    TRY(nf_get_node_field(s, NF_NODE_VZ, vz.data()));
    TRY(nf_get_node_field(s, NF_NODE_FR, fr.data()));
    TRY(nf_get_node_field(s, NF_NODE_FZ, fz.data()));
    TRY(nf_get_node_field(s, NF_NODE_ON_WALL, onw.data()));
    TRY(nf_get_node_field(s, NF_NODE_FIXED_R, fxr.data()));
    TRY(nf_get_node_field(s, NF_NODE_FIXED_Z, fxz.data()));
    std::vector<double> sig(4 * (size_t)P), work(P), epsP(P), hgR(E), hgZ(E), hgK(E), gone(E);
    TRY(nf_get_point_field(s, NF_POINT_STRESS, sig.data()));
    TRY(nf_get_point_field(s, NF_POINT_PLASTIC_WORK, work.data()));
    TRY(nf_get_point_field(s, NF_POINT_PLASTIC_STRAIN, epsP.data()));
    TRY(nf_get_element_field(s, NF_ELEMENT_HG_R, hgR.data()));
    TRY(nf_get_element_field(s, NF_ELEMENT_HG_Z, hgZ.data()));
    TRY(nf_get_element_field(s, NF_ELEMENT_HG_STIFFNESS, hgK.data()));
    TRY(nf_get_element_field(s, NF_ELEMENT_ERODED, gone.data()));
    Step h;
    TRY(s->readStep(h));

    double p = 0.0;
    for (int i = 0; i < N; i++) p += mass[i] * vz[i];
    o->axial_momentum = p;
    p = 0.0;
    for (int i = 0; i < N; i++) p += mass[i] * vr[i];
    o->radial_momentum = p;
    double ke = 0.0;
    for (int i = 0; i < N; i++) ke += mass[i] * (vr[i] * vr[i] + vz[i] * vz[i]);
    o->kinetic_energy = 0.5 * ke;
    ke = 0.0;
    const double hdt = 0.5 * h.dt;
    for (int i = 0; i < N; i++) {
        const double invM = 1.0 / mass[i];
        const bool held = s->hasWall && onw[i] != 0.0 && fz[i] <= 0.0;
        const double cvr = fxr[i] != 0.0 ? vr[i] : vr[i] + hdt * fr[i] * invM;
        const double cvz = fxz[i] != 0.0 || held ? vz[i] : vz[i] + hdt * fz[i] * invM;
        ke += 0.5 * mass[i] * (cvr * cvr + cvz * cvz);
    }
    o->centred_kinetic_energy = ke;

    auto matOf = [&](int e) -> const HostMaterial& { return s->materials[s->materialOf.empty() ? 0 : s->materialOf[e]]; };
    auto gaussVolume = [&](int e, int g) {
        const int* c = &s->conn[4 * e];
        const double r0 = s->r[c[0]], r1 = s->r[c[1]], r2 = s->r[c[2]], r3 = s->r[c[3]];
        const double z0 = s->z[c[0]], z1 = s->z[c[1]], z2 = s->z[c[2]], z3 = s->z[c[3]];
        const double G = 0.5773502691896258;
        const double GX[4] = {-G, G, G, -G}, GE[4] = {-G, -G, G, G};
        const double xi = ppe == 1 ? 0.0 : GX[g];
        const double eta = ppe == 1 ? 0.0 : GE[g];
        const double weight = ppe == 1 ? 4.0 : 1.0;
        const double xm = 1.0 - xi, xp = 1.0 + xi, em = 1.0 - eta, ep = 1.0 + eta;
        const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em, N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
        const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
        const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
        const double J00 = dN0x * r0 + dN1x * r1 + dN2x * r2 + dN3x * r3;
        const double J01 = dN0x * z0 + dN1x * z1 + dN2x * z2 + dN3x * z3;
        const double J10 = dN0e * r0 + dN1e * r1 + dN2e * r2 + dN3e * r3;
        const double J11 = dN0e * z0 + dN1e * z1 + dN2e * z2 + dN3e * z3;
        const double det = J00 * J11 - J01 * J10;
        const double rg = N0 * r0 + N1 * r1 + N2 * r2 + N3 * r3;
        return weight * (s->formulation == NF_AXISYMMETRIC ? 2.0 * 3.141592653589793 * rg * det : s->thickness * det);
    };
    double se = 0.0;
    for (int e = 0; e < E; e++) {
        if (gone[e] != 0.0) continue;
        const HostMaterial& m = matOf(e);
        const double lam = s->formulation == NF_PLANE_STRESS ? 2.0 * m.lambda() * m.mu() / (m.lambda() + 2.0 * m.mu()) : m.lambda();
        const double bulk = lam + 2.0 * m.mu() / 3.0;
        for (int g = 0; g < ppe; g++) {
            const size_t q = 4 * (size_t)(e * ppe + g);
            const double press = (sig[q] + sig[q + 1] + sig[q + 2]) / 3.0;
            const double a = sig[q] - press, bb = sig[q + 1] - press, c = sig[q + 2] - press, dd = sig[q + 3];
            const double density = (a * a + bb * bb + c * c + 2.0 * dd * dd) / (4.0 * m.mu()) + press * press / (2.0 * bulk);
            se += density * gaussVolume(e, g);
        }
    }
    o->strain_energy = se;
    double w = 0.0;
    for (int e = 0; e < E; e++)
        for (int g = 0; g < ppe; g++) w += work[e * ppe + g] * gaussVolume(e, g);
    o->plastic_dissipation = w;
    double hg = 0.0;
    if (!(s->integration == NF_FULL || s->d.qhg == 0.0))
        for (int e = 0; e < E; e++)
            if (hgK[e] > 0.0) hg += 0.5 * (hgR[e] * hgR[e] + hgZ[e] * hgZ[e]) / hgK[e];
    o->hourglass_energy = hg;
    double mx = 0.0;
    int yielded = 0;
    for (int q = 0; q < P; q++) {
        if (epsP[q] > mx) mx = epsP[q];
        if (epsP[q] > 0.0) yielded++;
    }
    o->max_plastic_strain = mx;
    o->yielded_fraction = (double)yielded / P;
    if (s->anyJC) {
        double t = -HUGE_VAL;
        bool ok = true;
        for (int q = 0; q < P && ok; q++) {
            const HostMaterial& m = matOf(q / ppe);
            if (!m.raw.has_johnson_cook) {
                ok = false;
                break;
            }
            const nf_johnson_cook& j = m.raw.johnson_cook;
            const double temp = j.room_temperature + j.taylor_quinney * work[q] / (m.raw.density * j.specific_heat);
            if (temp != temp) t = temp;
            else if (t == t && temp > t) t = temp;
        }
        o->max_temperature = ok ? t : NAN;
    } else {
        o->max_temperature = NAN;
    }
    if (s->hasDamage) {
        std::vector<double> st(P), so(P), fs(P), gf(P);
        TRY(nf_get_point_field(s, NF_POINT_DAMAGE_STRENGTH, st.data()));
        TRY(nf_get_point_field(s, NF_POINT_DAMAGE_SOFTENING, so.data()));
        HIP_TRY(hipMemcpyAsync(fs.data(), s->d.dFailure, sizeof(double) * P, hipMemcpyDeviceToHost, s->stream));
        TRY(s->sync());
        double fd = 0.0, md = 0.0;
        int begun = 0, failed = 0;
        for (int e = 0; e < E; e++)
            for (int g = 0; g < ppe; g++) {
                const int q = e * ppe + g;
                const double sf = st[q];
                double diss = 0.0;
                if (sf > 0.0) {  // Damage.dissipation :363-372
                    const double slope = so[q];
                    const double full = -sf / slope;
                    double excess = epsP[q] - fs[q];
                    if (excess > 0.0) {
                        if (excess > full) excess = full;
                        diss = excess * (sf + 0.5 * slope * excess);
                    }
                }
                fd += diss * gaussVolume(e, g);
            }
        for (int q = 0; q < P; q++) {
            double dmg = 0.0;
            if (st[q] > 0.0) {
                const double dd = -so[q] * (epsP[q] - fs[q]) / st[q];
                dmg = dd <= 0.0 ? 0.0 : (dd >= 1.0 ? 1.0 : dd);
            }
            if (dmg > md) md = dmg;
            if (st[q] > 0.0) begun++;
            if (dmg >= 1.0) failed++;
        }
        o->fracture_dissipation = fd;
        o->max_damage = md;
        o->damaged_fraction = (double)begun / P;
        o->failed_fraction = (double)failed / P;
    }
    return NF_OK;
}

// Per-kernel timing of the step loop: when enabled, every launch of the nodal, element,
// contact and fused step kernels is bracketed by HIP events on the handle's stream.
// nf_profile_read returns the summed milliseconds and launch counts (indices 0 nodal,
// 1 element, 2 contact, 3 fused step) and resets them.
int nf_profile(nf_solver* s, int32_t enable) {
    s->touch();
    g_error.clear();
    TRY(s->harvest());
    s->profiling = enable != 0;
    return NF_OK;
}

int nf_profile_read(nf_solver* s, double* ms, int64_t* counts) {
    s->touch();
    g_error.clear();
    TRY(s->harvest());
    for (int k = 0; k < 4; k++) {
        ms[k] = s->kernelMs[k];
        counts[k] = s->kernelCount[k];
        s->kernelMs[k] = 0.0;
        s->kernelCount[k] = 0;
    }
    return NF_OK;
}

// Drives `paths` independent material points of one material through `steps` strain
// increments each (J2.update, solver/J2.java:163-296).
int nf_point_path(const nf_material* m, int32_t formulation, int32_t paths, int32_t steps,
                  const double* increments, const double* damage, double* out) {
    g_error.clear();
    TRY(check_material(m));
    if (paths < 1 || steps < 1 || !increments || !out) return fail(NF_E_ARGUMENT, "empty path");
    const Mat mat = make_mat(*m, formulation);
    double *di = nullptr, *dd = nullptr, *dout = nullptr;
    const size_t ni = 5 * (size_t)paths * steps, no = 14 * (size_t)paths * steps;
    TRY(dev_alloc(&di, ni));
    TRY(dev_alloc(&dout, no));
    if (damage) TRY(dev_alloc(&dd, 7 * (size_t)paths));
    HIP_TRY(hipMemcpy(di, increments, sizeof(double) * ni, hipMemcpyHostToDevice));
    if (damage) HIP_TRY(hipMemcpy(dd, damage, sizeof(double) * 7 * paths, hipMemcpyHostToDevice));
    if (damage) hipLaunchKernelGGL(k_point_path<true>, dim3(blocks_for(paths)), dim3(BLOCK), 0, 0, mat, paths, steps, (const double*)di, (const double*)dd, dout);
    else hipLaunchKernelGGL(k_point_path<false>, dim3(blocks_for(paths)), dim3(BLOCK), 0, 0, mat, paths, steps, (const double*)di, (const double*)dd, dout);
    hipError_t err = hipMemcpy(out, dout, sizeof(double) * no, hipMemcpyDeviceToHost);
    (void)hipFree(di);
    (void)hipFree(dd);
    (void)hipFree(dout);
    if (err != hipSuccess) return fail(NF_E_DEVICE, hipGetErrorString(err));
    return NF_OK;
}

// Device-to-device copy bandwidth: `bytes` copied `reps` times by k_copy; returns read plus
// written bytes per second in GB/s and the mean kernel time.
int nf_copy_bandwidth(int64_t bytes, int32_t reps, double* gbps, double* ms_per_copy) {
    g_error.clear();
    if (bytes < 16 || reps < 1) return fail(NF_E_ARGUMENT, "need at least 16 bytes and one repetition");
    const size_t n = (size_t)bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    TRY(dev_alloc(&a, n));
    if (int rc = dev_alloc(&b, n)) {
        (void)hipFree(a);
        return rc;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int grid = 256 * 16;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(BLOCK), 0, 0, (const double2*)a, b, n);
    HIP_TRY(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(BLOCK), 0, 0, (const double2*)a, b, n);
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    *ms_per_copy = ms / reps;
    *gbps = 2.0 * 16.0 * n / (*ms_per_copy * 1e-3) / 1e9;
    return NF_OK;
}

// Evaluates the device implementations of the shared transcendental functions (and the
// device's division and square root) on n inputs: a self-check that the GPU computes exactly
// what the host computes. fn: 0 log, 1 exp, 2 pow, 3 atan2(x, y), 4 sin, 5 cos, 6 hypot,
// 7 sqrt, 8 x / y, 9 log1p, ... (include/neofiz.h lists them all). For the solve (21-23) x
// holds 5 n values and y 5.
int nf_math_eval(int32_t fn, const double* x, const double* y, double* out, int32_t n) {
    g_error.clear();
    if (n <= 0) return NF_OK;
    if (fn < 0 || fn > 25) return fail(NF_E_ARGUMENT, "unknown function");
    const bool solve = fn >= 21 && fn <= 23;
    if (solve && !y) return fail(NF_E_ARGUMENT, "the solve needs the material values");
    const size_t nx = solve ? 5 * (size_t)n : (size_t)n, ny = solve ? 5 : (size_t)n;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    TRY(dev_alloc(&dx, nx));
    TRY(dev_alloc(&dy, ny));
    TRY(dev_alloc(&dout, n));
    HIP_TRY(hipMemcpy(dx, x, sizeof(double) * nx, hipMemcpyHostToDevice));
    if (y) HIP_TRY(hipMemcpy(dy, y, sizeof(double) * ny, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_math_eval, dim3(blocks_for(n)), dim3(BLOCK), 0, 0, fn, (const double*)dx,
                       (const double*)dy, dout, n);
    hipError_t err = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    if (err != hipSuccess) return fail(NF_E_DEVICE, hipGetErrorString(err));
    return NF_OK;
}


// ------------------------------------------------------------------ sensitivities

int nf_enable_sensitivity(nf_solver* s) {
    s->touch();
    g_error.clear();
    if (s->tangent) return NF_OK;
    Step h;
    TRY(s->readStep(h));
    if (h.steps != 0) return fail(NF_E_STATE, "sensitivities start from the initial state; enable them before the first step");
    if (s->adjoint) return fail(NF_E_STATE, "a solver carries forward sensitivities or the adjoint, not both");
    if (s->integration != NF_REDUCED || s->hasDamage || s->hasErosion || s->hasContact || s->edges ||
        !s->materialOf.empty() || !s->anyJC || s->hasKin)
        return fail(NF_E_STATE, "sensitivities cover reduced integration of one Johnson-Cook material without "
                                "kinematic hardening, damage, erosion, contact or pressure edges");
    TDev& t = s->td;
    const size_t N = s->N, E = s->E, P = s->P;
    TRY(s->alloc(&t.tu, NP * N));
    TRY(s->alloc(&t.tv, NP * N));
    TRY(s->alloc(&t.tf, NP * N));
    TRY(s->alloc(&t.tslot, NP * 4 * E));
    TRY(s->alloc(&t.tsig, NP * 4 * P));
    TRY(s->alloc(&t.tep, NP * P));
    TRY(s->alloc(&t.twork, NP * P));
    TRY(s->alloc(&t.thg, NP * 2 * E));
    TRY(s->alloc(&t.tedge, NP * E));
    TRY(s->alloc(&t.edgeVal, E));
    TRY(s->alloc(&t.tstep, 4 * NP));
    TRY(s->alloc(&t.argmin, 3));
    const int none[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0};
    HIP_TRY(hipMemcpy(t.argmin, none, sizeof none, hipMemcpyHostToDevice));
    s->tangent = true;
    s->graphDirty = true;
    return NF_OK;
}

int nf_get_node_sensitivity(nf_solver* s, int32_t field, int32_t param, double* out) {
    s->touch();
    g_error.clear();
    if (!s->tangent) return fail(NF_E_STATE, "sensitivities are not enabled");
    if (param < 0 || param >= NP) return fail(NF_E_ARGUMENT, "parameter out of range");
    const double2* src = nullptr;
    switch (field) {
        case NF_NODE_UR: case NF_NODE_UZ: src = s->td.tu; break;
        case NF_NODE_VR: case NF_NODE_VZ: src = s->td.tv; break;
        case NF_NODE_FR: case NF_NODE_FZ: src = s->td.tf; break;
        default: return fail(NF_E_ARGUMENT, "sensitivities exist for UR, UZ, VR, VZ, FR and FZ");
    }
    std::vector<double2> h(s->N);
    HIP_TRY(hipMemcpyAsync(h.data(), src + (size_t)param * s->N, sizeof(double2) * s->N, hipMemcpyDeviceToHost, s->stream));
    TRY(s->sync());
    const bool r = field == NF_NODE_UR || field == NF_NODE_VR || field == NF_NODE_FR;
    for (int i = 0; i < s->N; i++) out[i] = r ? h[i].x : h[i].y;
    return NF_OK;
}

int nf_get_step_sensitivity(nf_solver* s, double* out) {
    s->touch();
    g_error.clear();
    if (!s->tangent) return fail(NF_E_STATE, "sensitivities are not enabled");
    TRY(s->sync());
    HIP_TRY(hipMemcpy(out, s->td.tstep + (size_t)s->q * 2 * NP, sizeof(double) * 2 * NP, hipMemcpyDeviceToHost));
    return NF_OK;
}

int nf_adjoint_enable(nf_solver* s, int32_t checkpoint_interval) {
    s->touch();
    g_error.clear();
    if (s->adjoint) return NF_OK;
    if (checkpoint_interval < 1 || checkpoint_interval > 4096)
        return fail(NF_E_ARGUMENT, "the checkpoint interval lies in [1, 4096] steps");
    Step h;
    TRY(s->readStep(h));
    if (h.steps != 0) return fail(NF_E_STATE, "the adjoint records from the initial state; enable it before the first step");
    if (s->tangent) return fail(NF_E_STATE, "a solver carries forward sensitivities or the adjoint, not both");
    if (s->integration != NF_REDUCED || s->hasDamage || s->hasErosion || s->hasContact || s->edges ||
        !s->materialOf.empty() || !s->anyJC || s->hasKin)
        return fail(NF_E_STATE, "the adjoint covers reduced integration of one Johnson-Cook material without "
                                "kinematic hardening, damage, erosion, contact or pressure edges");
    s->adjK = checkpoint_interval;
    s->adjoint = true;
    s->graphDirty = true;
    return NF_OK;
}

int nf_adjoint_gradient(nf_solver* s, int32_t objective_count, const double* seeds, double* gradient) {
    s->touch();
    g_error.clear();
    if (!s->adjoint) return fail(NF_E_STATE, "the adjoint is not enabled");
    if (objective_count < 1 || objective_count > AW_MAX)
        return fail(NF_E_ARGUMENT, "one or two objectives per reverse sweep");
    if (!seeds || !gradient) return fail(NF_E_ARGUMENT, "null array");
    if (s->checkpoints.empty() || s->finalState.steps == 0) return fail(NF_E_STATE, "no steps were recorded");
    Step h;
    TRY(s->readStep(h));
    if (h.steps != s->finalState.steps) return fail(NF_E_STATE, "the solver moved past the recorded run");
    return s->adjointSweep(objective_count, seeds, gradient);
}

int nf_adjoint_info(nf_solver* s, int64_t* out) {
    s->touch();
    g_error.clear();
    out[0] = s->finalState.steps;
    out[1] = (int64_t)s->checkpoints.size();
    out[2] = s->adjCaps;
    out[3] = s->adjK;
    return NF_OK;
}

// ------------------------------------------------------------------ batched ensemble

int nf_batch_create(nf_batch** out, int32_t count, nf_solver* const* solvers) {
    g_error.clear();
    if (!out) return fail(NF_E_ARGUMENT, "out must not be null");
    *out = nullptr;
    if (count < 1 || !solvers) return fail(NF_E_ARGUMENT, "a batch needs at least one solver");
    int integration = -1;
    for (int i = 0; i < count; i++) {
        nf_solver* s = solvers[i];
        if (!s) return fail(NF_E_ARGUMENT, "null solver in batch");
        for (int j = 0; j < i; j++)
            if (solvers[j] == s) return fail(NF_E_ARGUMENT, "a solver appears twice in the batch");
        if (s->hasContact || s->hasDamage || s->hasErosion || s->edges || !s->materialOf.empty())
            return fail(NF_E_STATE, "a batch carries neither contact, damage, erosion, pressure edges nor "
                                    "several materials in one instance");
        if (integration < 0) integration = s->integration;
        if (s->integration != integration) return fail(NF_E_ARGUMENT, "every instance of a batch uses the same integration");
        TRY(s->sync());
    }
    nf_batch* b = new nf_batch();
    auto bail = [&](int rc) {
        delete b;
        return rc;
    };
    int rc;
    if (const char* g = std::getenv("NEOFIZ_BATCH_GRAPHS")) b->useGraphs = g[0] != '0';
    if (const char* g = std::getenv("NEOFIZ_BATCH_PERSISTENT")) b->persistent = g[0] == '1';
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NF_E_DEVICE, "hipStreamCreate failed: no usable GPU"));
    b->full = integration == NF_FULL;
    const int ppe = b->full ? 4 : 1;
    std::vector<int4> conn;
    std::vector<int> slotStart(1, 0), slotOf;
    std::vector<int4> slot4;
    for (int i = 0; i < count; i++) {
        nf_solver* s = solvers[i];
        b->solvers.push_back(s);
        BInst I;
        std::memset(&I, 0, sizeof I);
        I.mat = s->d.mat0;
        I.wall = s->hasWall ? 1 : 0;
        I.wallFrictional = s->d.wallFrictional;
        I.friction = s->d.friction;
        I.plane = s->hasWall ? s->d.plane : -HUGE_VAL;  // no wall: no node ever arrives
        I.cfl = s->d.cfl;
        I.fastestWave = s->d.fastestWave;
        I.massScale = s->d.massScale;
        I.gravityR = s->d.gravityR;
        I.gravityZ = s->d.gravityZ;
        I.damping = s->d.damping;
        I.qhg = s->d.qhg;
        I.thickness = s->d.thickness;
        I.axi = s->d.axi;
        I.finite = s->d.finite;
        I.nodeBase = b->N;
        I.nodes = s->N;
        I.elemBase = b->E;
        I.elems = s->E;
        b->inst.push_back(I);
        b->instTs.push_back(BLOCK);
        b->live.push_back(1);
        if (s->hasKin) b->kin = true;
        if (s->hasWall) b->wall = true;
        for (int e = 0; e < s->E; e++)
            conn.push_back(make_int4(s->conn[4 * e] + b->N, s->conn[4 * e + 1] + b->N, s->conn[4 * e + 2] + b->N,
                                     s->conn[4 * e + 3] + b->N));
        const int slotBase = (int)slotOf.size();
        for (int n = 0; n < s->N; n++) {
            slotStart.push_back(slotBase + s->slotStart[n + 1]);
            int k[4] = {-1, -1, -1, -1};
            for (int c = 0; c < 4 && s->slotStart[n] + c < s->slotStart[n + 1]; c++)
                k[c] = s->slotOf[s->slotStart[n] + c] + 4 * b->E;
            slot4.push_back(make_int4(k[0], k[1], k[2], k[3]));
        }
        for (int x : s->slotOf) slotOf.push_back(x + 4 * b->E);
        b->N += s->N;
        b->E += s->E;
    }
    b->P = b->E * ppe;
    Dev& d = b->d;
    d.N = b->N;
    d.E = b->E;
    d.P = b->P;
    d.ppe = ppe;
    d.erosionDamage = HUGE_VAL;
    double2* x0 = nullptr;
    double* z0 = nullptr;
    double* mass = nullptr;
    int4* conn4 = nullptr;
    int* ss = nullptr;
    int* so = nullptr;
    int4* s4 = nullptr;
    const size_t tilesMax = (size_t)b->E / 64 + b->N / 64 + 4 * count + 8;
    if ((rc = b->alloc(&x0, b->N)) || (rc = b->alloc(&z0, b->N)) || (rc = b->alloc(&d.u, b->N)) ||
        (rc = b->alloc(&d.v, b->N)) || (rc = b->alloc(&d.f, b->N)) || (rc = b->alloc(&mass, b->N)) ||
        (rc = b->alloc(&d.flags, b->N)) || (rc = b->alloc(&conn4, b->E)) || (rc = b->alloc(&d.sig0, b->P)) ||
        (rc = b->alloc(&d.sig1, b->P)) || (rc = b->alloc(&d.sig2, b->P)) || (rc = b->alloc(&d.sig3, b->P)) ||
        (rc = b->alloc(&d.back0, b->P)) || (rc = b->alloc(&d.back1, b->P)) || (rc = b->alloc(&d.back2, b->P)) ||
        (rc = b->alloc(&d.back3, b->P)) || (rc = b->alloc(&d.epsP, b->P)) || (rc = b->alloc(&d.work, b->P)) ||
        (rc = b->alloc(&d.hgR, b->E)) || (rc = b->alloc(&d.hgZ, b->E)) || (rc = b->alloc(&d.hgK, b->E)) ||
        (rc = b->alloc(&d.slot, 4 * (size_t)b->E)) || (rc = b->alloc(&ss, b->N + 1)) ||
        (rc = b->alloc(&so, slotOf.size())) || (rc = b->alloc(&s4, b->N)) ||
        (rc = b->alloc(&b->instDev, count)) || (rc = b->alloc(&b->steps, 2 * (size_t)count)) ||
        (rc = b->alloc(&b->eBlocks, tilesMax)) || (rc = b->alloc(&b->eTiles, tilesMax)) ||
        (rc = b->alloc(&b->nBlocks, tilesMax)) || (rc = b->alloc(&b->nTiles, tilesMax)))
        return bail(rc);
    d.x0 = x0;
    d.z0 = z0;
    d.mass = mass;
    d.conn = conn4;
    d.slotStart = ss;
    d.slotOf = so;
    d.slot4 = s4;
    if (hipMemcpy(conn4, conn.data(), sizeof(int4) * conn.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ss, slotStart.data(), sizeof(int) * slotStart.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(so, slotOf.data(), sizeof(int) * slotOf.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s4, slot4.data(), sizeof(int4) * slot4.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->instDev, b->inst.data(), sizeof(BInst) * count, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(NF_E_DEVICE, "upload failed"));
    if ((rc = batch_transfer(b, true))) return bail(rc);
    *out = b;
    return NF_OK;
}

int nf_batch_destroy(nf_batch* b) {
    delete b;
    return NF_OK;
}

int nf_batch_set_tile(nf_batch* b, int32_t instance, int32_t threads) {
    g_error.clear();
    if (instance < -1 || instance >= (int)b->inst.size()) return fail(NF_E_ARGUMENT, "instance out of range");
    if (threads != 64 && threads != 128 && threads != 256) return fail(NF_E_ARGUMENT, "a tile is 64, 128 or 256 threads");
    for (int i = 0; i < (int)b->inst.size(); i++)
        if (instance < 0 || i == instance) b->instTs[i] = threads;
    b->graphDirty = true;
    return NF_OK;
}

int nf_batch_get_tile(nf_batch* b, int32_t instance, int32_t* threads) {
    g_error.clear();
    if (instance < 0 || instance >= (int)b->inst.size() || !threads) return fail(NF_E_ARGUMENT, "instance out of range");
    *threads = b->instTs[instance];
    return NF_OK;
}

int nf_batch_run_until(nf_batch* b, double end_time, int64_t max_steps) {
    g_error.clear();
    return b->runUntil(end_time, max_steps > 0 ? max_steps : -1);
}

int nf_batch_scatter(nf_batch* b) {
    g_error.clear();
    return batch_transfer(b, false);
}

}  // extern "C"
