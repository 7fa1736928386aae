// Device-side data of one sweep handle and the launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace allium {

// Everything a sweep kernel reads, passed by value.
struct SweepParams {
    int dim, p, n, nf;
    int Ne, Nomega, G;
    int nel[3];
    int stride[3];          // element id strides per axis
    float inv_nel0, inv_nel01;
    long long Ndof;
    const double *omega;    // Nomega x 4 (x, y, z, w)
    // geometry: uniform => one element's matrices; general => per element / face
    int uniform;
    const double *K;        // uniform: Nomega x n x n, streaming + outflow faces + tau * Mt
    const float *Kf;
    const double *Mu;       // unit mass (MG rule): uniform n x n, general Ne x n x n
    const float *Muf;
    const double *Mt;       // time mass (= MassIntegrator rule matrix)
    const float *Mtf;
    const double *Gd;       // general: Ne x dim x n x n
    const double *Fface;    // uniform: 2dim x nf x nf; general: Nf x nf x nf
    const float *Ffacef;
    const double *fnor;     // general: Nf x 3
    int foff[4];
    int fext[3][3];
    const double *sigma;    // G x Ne
    double tau;
    int timedep;
    const int *refl;        // Nomega x 2dim reflected direction or -1
    int refl_mask;          // bit s set => side s reflective
    // vectors
    const double *src;      // psi-sized source (may be null)
    const double *iso;      // phi-sized isotropic source added to every direction (may be null)
    double *psi;
    const double *zeros;    // 64 zero doubles: the target of loads a kernel issues unconditionally
    const uint32_t *cells;  // (e << 10) | a, grouped by level
    // negative-flux fixup after each cell solve (src/fixup.cpp): 0 off,
    // 1 zero, 2 zero and scale, 3 ryosuke
    int fixup;
    double fix_min;
    unsigned long long *fix_counts;   // applied, failed, ryosuke beyond 3 positive unknowns
    double *fix_monitor;              // G x Ne application counts, or null
    // 1: the box sweep writes psi with nontemporal stores (psi is read once,
    // by the moment kernel, before the next sweep)
    int nt_psi;
};

// One level of the direction-uniform wave plan: offset of the level in the
// padded cell array and the prefix of 64-cell waves per XCD bin.
struct WaveLevel {
    long long base;   // offset of the level in the padded cell array
    int nparts;       // octant groups in the plan (1)
    int pb[33];       // prefix of waves over (group, bin)
};

// Tiles of the persistent kernel: a run of (cell, group) items of one
// (level, direction) class.
struct Tile {
    int key;        // dependency class of this tile
    int wait_key;   // class that must be complete first, -1 if none
    int cell_begin; // first cell of the class in the cells array
    int ncells;     // cells in the class
    int item_begin; // first (cell, group) item of this tile
    int nitems;
};

// Box sweeps (uniform geometry, directions that do not couple).  The mesh is
// cut into boxes of bs[0] x bs[1] x bs[2] elements; a workgroup sweeps one box
// for nd directions of one octant and one group, passing the face traces
// between the box's cells through LDS planes.  Boxes on one anti-diagonal of
// the box grid (in the octant's orientation) are independent: one launch per
// box level.
// doubles per trace slot in the box sweep's LDS planes: an odd stride so that
// the slots lanes read at once spread over the LDS banks
__host__ __device__ constexpr int box_slot_stride(int nf) { return nf % 2 == 0 ? nf + 1 : nf; }

struct BoxPlan {
    int bs[3];               // box extents in elements (1 beyond dim)
    int nb[3];               // boxes per axis
    int nd;                  // directions per workgroup
    int nchunk;              // workgroups per (box, octant, group)
    int maxoct;              // directions per octant row of oct_dirs
    int noct;                // 2^dim
    int nl;                  // levels inside a full box
    int kstride;             // doubles per direction's matrix in LDS
    int psz[3];              // LDS plane slots per axis (product of the other extents)
    int lds_bytes;           // planes, matrices, |Omega|, directions, cells, level offsets
    double vol;              // element volume: Mu = vol * (tensor product of 1D unit masses)
    double farea[3];         // area of the element faces normal to each axis
    const int *oct_dirs;     // noct x maxoct direction ids, -1 padded
    const uint32_t *lcells;  // cells of a full box by level: li | lj << 10 | lk << 20
    const int *lptr;         // nl + 1
    const uint32_t *boxes;   // oriented box coordinates by box level: bi | bj << 10 | bk << 20
};

struct DeviceSweep {
    SweepParams prm{};
    int precision = 0;       // 0 fp64 cell solves, 1 fp32 cell solves
    // fixup state kept while it is switched off (UseFixup(false))
    int fix_type = 0;
    bool fix_on = false;
    int launch_mode = 0;     // 0 levels, 1 graph, 2 persistent, 3 direction sweeps, 4 box sweeps,
                             // 5 box sweeps in a graph
    int dir_block = 512;     // workgroup size of the direction-sweep kernel
    int block = 128;
    int lanes = 4;           // lanes per cell in the level kernel (1, 2 or 4); 0 = direction-uniform
                             // waves over XCD bins (uniform geometry)
    int persistent_tile = 256;
    // box sweeps (launch modes 4 and 5)
    BoxPlan box{};
    std::vector<int> box_ptr;    // box level -> first entry of box.boxes
    int box_threads = 256;
    int box_ncu = 0;             // CUs of the device (box launches of at most one workgroup per CU)
    int refine = 0;              // box sweeps at p = 1: in-place defect correction with FP32 cell solves
    int stream_psi = 0;          // set by the solvers whose every sweep is followed by the moment kernel
    int box_prefetch = 2;        // 0 none, 2 next level's inputs loaded during the current solve
    void *box_buf = nullptr;     // device memory of the box plan tables
    // host copies of the plan
    std::vector<long> level_ptr;
    int nlevels = 0;
    // device buffers owned by the handle
    std::vector<void *> owned;
    double *d_omega = nullptr;
    uint32_t *d_cells = nullptr;
    uint32_t *d_wcells = nullptr;
    long long *d_cls_ptr = nullptr;  // nlevels x Nomega + 1: start of each (level, direction) class
    bool angle_independent = true;
    std::vector<long> wlevel_base;
    std::vector<int> wpb;
    int nparts = 1;

    int *d_refl = nullptr;
    // persistent kernel state
    Tile *d_tiles = nullptr;
    int ntiles = 0;
    int nkeys = 0;
    int *d_counters = nullptr;  // nkeys + 2 (ticket, error flag)
    int *d_need = nullptr;
    // graph cache
    hipGraphExec_t graph_exec = nullptr;
    const void *graph_src = nullptr, *graph_iso = nullptr;
    void *graph_psi = nullptr;
    int graph_prec = -1;
    hipStream_t graph_stream = nullptr;
    // timing (TimingLog "sweep", EventLog "sweeps"): launch_sweep records ev0 /
    // ev1 around each sweep.  The caller points them at a ring of event pairs
    // and reads a pair only when it comes round again or the total is asked
    // for, so timing never makes the host wait for a sweep.
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kTimingRing = 64;
    hipEvent_t ring0[kTimingRing] = {}, ring1[kTimingRing] = {};
    bool ring_busy[kTimingRing] = {};
    int ring_next = 0;
    double sweep_seconds = 0.0;
    long sweeps = 0;
    bool timing = true;
};

// the level kernel's negative-flux fixups exist for these
bool fixup_supported(int dim, int p);
// psi = L^{-1}(src + iso); with phi, the moment kernel then forms phi = D psi
hipError_t launch_sweep(DeviceSweep &ds, const double *src, const double *iso, double *psi,
                        hipStream_t stream, double *phi = nullptr);

// Cooperative cell solves (local systems too large for one lane: 3D p = 2):
// 16 lanes per cell, lane t owns rows t and t + 16, kCoopCells cells per
// workgroup.  One launch runs one level of up to kMaxBatch independent
// instances (batched sweeps of uneven meshes).
constexpr int kMaxBatch = 16;
constexpr int kCoopCells = 16;
struct CoopLevel {
    int ninst;
    int start[kMaxBatch + 1];    // prefix over instances of (cells x groups) in this launch
    long long cbeg[kMaxBatch];   // first cell of the level in each instance's cells array
    int ncells[kMaxBatch];       // cells of the level in each instance
};
bool coop_supported(int dim, int p);
// d_prms: device array of the instances' parameters, or null with ninst == 1 to use p0
hipError_t launch_coop_level(const SweepParams &p0, const SweepParams *d_prms, const CoopLevel &cl, int dim,
                             int p, bool uniform, int precision, hipStream_t s);
hipError_t reset_persistent(DeviceSweep &ds, hipStream_t stream);
// one sweep as box-level launches (modes 4/5); the box plan must be built
hipError_t launch_box_sweep(const DeviceSweep &ds, const SweepParams &prm, hipStream_t stream);
bool box_sweep_supported(int dim, int p);
// entry (i, j) of the unit-volume tensor-product mass the box kernel uses
double host_unit_mass(int dim, int p, int i, int j);
bool sweep_supported(int dim, int p, std::string &why);

// streaming operator L psi (AdvectionOperator::Mult)
hipError_t launch_apply_L(const SweepParams &prm, const double *x, double *y, hipStream_t stream);
// y = q - L x (the FP64 defect of the FP32-refined source iteration); y may not alias x
hipError_t launch_defect_L(const SweepParams &prm, const double *q, const double *x, double *y, hipStream_t s);
// per-block partial sums of absorption and leakage (ComputeBalance); out has
// 2 * balance_blocks(prm) doubles
int balance_blocks(const SweepParams &prm);
hipError_t launch_balance(const SweepParams &prm, const double *psi, double *partial,
                          hipStream_t stream);
// the balance without allocation: partials and result in the file-scope device
// workspace of the calling GPU (balance_blocks(prm) <= kStaticBalanceBlocks,
// else hipErrorInvalidValue); synchronizes the stream.  Calls that overlap on
// one GPU share the workspace and race.
constexpr int kStaticBalanceBlocks = 32768;   // Ne * G <= 8,388,608
constexpr int kMaxDevices = 64;
hipError_t balance_static(const SweepParams &prm, const double *psi, double *absorption, double *leakage,
                          hipStream_t stream);
// phi = sum_a w_a psi ; psi = phi / wsum ; y = scale * coef_e * Mm(e) x_e
hipError_t launch_moments(const SweepParams &prm, const double *psi, double *phi, hipStream_t s);
// [phi, J] ((dim + 1) G Ndof, moment slowest) and its transpose
hipError_t launch_moments_aniso(const SweepParams &prm, int dim, const double *psi, double *mom, hipStream_t s);
hipError_t launch_moments_aniso_T(const SweepParams &prm, int dim, double wsum, const double *mom, double *psi,
                                  hipStream_t s);
hipError_t launch_moments_T(const SweepParams &prm, double inv_wsum, const double *phi, double *psi,
                            hipStream_t s);
hipError_t launch_mass(const SweepParams &prm, const double *Mm, const double *coef, double scale,
                       const double *x, double *y, hipStream_t s);
// SI update: r = b - (x - y) (y null: r = b - x); x += r; partial[blk] = sum r^2;
// rout (may be null) receives r
int si_update_blocks(long n);
hipError_t launch_si_update(long n, const double *b, const double *y, double *x, double *partial,
                            hipStream_t s, double *rout = nullptr);
// plain vector ops
hipError_t launch_axpby(long n, double a, const double *x, double b, const double *y, double *z,
                        hipStream_t s);

}  // namespace allium
