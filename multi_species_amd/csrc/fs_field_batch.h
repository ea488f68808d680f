// Batched periodic direct sum over independent point segments (decks).
// Internal to libfarrsight_hip.
#pragma once

#include <vector>

#include "fs_internal.h"
#include "fs_sym.h"

namespace fs {

enum { kPrecisionF64 = 0, kPrecisionMixed = 1 };

// Governor state per segment: max |E| and the error at each of 8 candidate
// FP64 bands (mixed precision).
constexpr int kGovVals = 9;

// Minimum FP64 band B0 of the mixed method for delta = eps/L (as a fraction
// of the period), and the governor's candidate band q = 0..7,
// B0 2^((q-2)/2), at least 2^-6 (>= 1/4: every pair in FP64).
double min_band(double delta);
double governor_band(double delta, int q);

// One segment of the concatenated point arrays. A deck is one segment that
// is its own source; the single-call API with distinct targets uses two
// segments (the sources, not evaluated, and the targets, sourced from the
// first).
struct BatchSeg {
    int off = 0, n = 0;  // points [off, off + n)
    int src = 0;         // segment whose points are the sources
    bool eval = true;    // evaluate the field at this segment's points
    bool open = false;   // open boundaries (FP64 only): E_MQ_DirectSum_openbcs
    double L = 1.0, eps = 0.0;
};

// Per-segment launch parameters (device copy).
struct SegDev {
    int sym;          // FP64: index of the segment's pair-once plan, or -1
    int off, n;
    int src_off, src_n;
    int splits, chunk;
    int eval, all_band;
    int open;         // open boundaries: z = x_i - x_j, delta2 = eps^2, E = sum - x_i
    int64_t part_base;
    double invL, delta2, C;
    // mixed only
    uint64_t band;
    double d2s;       // delta^2 2^128 (squared units of the fixed point)
    double b1z, b2z;  // series coefficients -3/4 and 5/8 in powers of delta^2 2^64
    int pw_base;
    int src;          // index of the source segment
    int is_src;       // some segment takes its sources from this one
    int all_band0;    // all_band of the minimum band (no governor)
    uint64_t band0;   // the minimum band B0 * 2^64
    uint64_t xband;   // far tiles whose sources lie within xband of their centre expand
    int tile_base;    // index of this segment's first source tile (source segments)
    int src_tile_base;
    int group_base;   // index of this segment's first group of 4 tiles
    int src_group_base;
    int log2h;        // expansion scale H = 2^log2h, the power of 2 at or below B0
    // far-field units of 512 and 2048 sources: index of this segment's first
    // record in the group arrays (after every group), and its source's
    int unit_base[2];
    int src_unit_base[2];
    int levels;       // unit levels the pair kernel uses for this segment's sources (0..2)
};

// Span of 4 consecutive tiles (128 sorted sources) and their largest emax:
// a group that is far and compact is summed without per-tile tests. The
// same record serves the far-field units of 512 and 2048 sources, with the
// unit's own emax (largest distance from its expansion's centre).
struct GroupRec {
    uint64_t vf, vl;  // first and last key, widened by the unsorted low bits
    uint64_t emax;
    uint64_t pad;
};

// Far-field expansion of one 32-source tile (mixed precision): with Z the
// displacement from the tile's centre to a target and rho = H/Z, the tile's
// regularization remainder at the target is rho |rho| sum_m c[m] rho^m.
struct TileRec {
    uint32_t hc;      // high word of the centre
    uint32_t pad;
    uint64_t emax;    // largest |centre - source| (units of 2^-64)
    float c[12];
};

// The accuracy probe's FP64 reference for a compact far tile: the same
// expansion carried to 7 terms of the series in d^2/z^2 and 12 moments (so
// that it sees the truncations of the FP32 expansion as well as its
// rounding), and the tile's moments sum w and sum w e/H.
struct TileRec64 {
    double c[25];
    double m0, m1;
};

// Moments M_k = sum_j w_j e_j^k (e = centre - source) of one tile, for the
// FP64 near-band expansion of a compact tile on one side of its targets,
// sum_j w_j z_j / sqrt(z_j^2 + d^2) = sum_k P_k(-Z/r) r^-(k+1) (Z M_k + M_k+1)
// with r = sqrt(Z^2 + d^2) (the Legendre generating function), k < 12,
// stored scaled by gamma_k = C(2k, k) / 2^k: a_k = gamma_k M_k and
// b_k = gamma_k M_k+1.
struct TileMom {
    double a[12];
    double b[12];
};

// Far-field record of one sorted source (FP32 path of the mixed kernel).
struct FarSrc {
    uint32_t h;             // high word of the position
    float w, wb1, wb2;      // -w delta^2/2 and its products with b1z, b2z
};

struct WorkItem {
    int seg;
    int tb, te;   // targets [tb, te) (concatenated index; sorted position when mixed)
    int sb, se;   // sources [sb, se)
    int split;
};

struct ScanBlock {
    int seg;
    int b, e;         // sorted positions [b, e)
    int first;        // index of the segment's first scan block
    int count;        // scan blocks of the segment
};

// Pair-once plan of one FP64 segment (its own source; fs_sym.h).
// A batch runs each such segment in one pass (sym::pass(P, 0) covers every
// chunk); segments whose single-pass partials exceed sym::kBatchDeckLimit
// take the full sum instead, a choice from the segment's size alone.
struct SymSeg {
    sym::Plan P;
    sym::Pass S;
    int seg, off;
    int item_base;          // first workgroup of the segment in its launch
    int64_t rbase, cbase;   // row and column partials (doubles into the work space)
    double invL, delta2;
    int open;               // open-boundary kernel (delta2 = eps^2)
};

// Plan and device tables of one batched field evaluation; the tables live
// in the object and are rebuilt by plan(). run() only launches kernels on
// the given stream (no allocation, no memset, no host synchronization), so
// a caller may capture it in a HIP graph.
class FieldBatch {
public:
    int plan(const std::vector<BatchSeg>& segs, int n_total, int precision, int num_cus,
             const FieldTuning& tun, hipStream_t stream);
    // Accuracy governor (mixed precision; off after plan()). When on, each
    // run first measures on its own inputs the E error that each candidate
    // FP64 band would give (probed targets against the exact FP64 field) and
    // takes the narrowest band within target * max|E| (target <= 0: 4e-9).
    void govern(bool on, double target = 0);
    // Governed runs: whether the next runs measure (probe and governor) or
    // keep every segment's band from the last run that did (on by default).
    void set_probe(bool on) { probe_ = on; }
    // The last governed run's measurements (nseg * kGovVals values: per
    // segment max|E| and the error at each candidate band, as doubles).
    int get_gov(std::vector<double>& out, hipStream_t stream) const;
    // Launch shape of the pair kernel: targets per thread (1, 2, 4; always 4
    // in mixed precision) and persistent workgroups per CU. Does not change
    // any result.
    int set_shape(int tpt, int wgs_per_cu, hipStream_t stream);
    // Times every shape on this batch's data, keeps the fastest (es is used
    // as scratch). Synchronizes the stream.
    int autotune(const double* x, const double* w, double* es, hipStream_t stream,
                 int* best_tpt, int* best_wgs);
    // x, w: concatenated positions and weights; es: field at every point of
    // the evaluated segments (others untouched).
    int run(const double* x, const double* w, double* es, hipStream_t stream);
    void release();
    // buffers from (and, at release, back to) a block pool when every use
    // of this batch is on the pool's stream (a simulation's)
    void set_pool(BlockPool* p) { pool_ = p; }
    int precision() const { return precision_; }
    int tpt() const { return tpt_; }
    int wgs_per_cu() const { return wgs_per_cu_; }
    int items() const { return (int)items_.size(); }
    int64_t pairs() const { return pairs_; }
    // FP64 kernel evaluations of one run (the pair-once segments evaluate
    // each pair once)
    int64_t evals() const { return evals_; }

private:
    void build_items(int tpt, std::vector<WorkItem>& items) const;
    int upload_items(const std::vector<WorkItem>& items, hipStream_t stream);
    int launch_pairs(int nit, int grid, hipStream_t stream);

    int precision_ = kPrecisionF64;
    bool governed_ = false;
    bool probe_ = true;
    double gov_target_ = 4e-9;
    uint64_t* gov() const;
    int n_total_ = 0, nseg_ = 0, tpt_ = 2, grid_ = 0, wgs_per_cu_ = 8, num_cus_ = 256;
    const double* x_ = nullptr;  // inputs of the current run (FP64 pair kernel)
    const double* w_ = nullptr;
    int64_t pairs_ = 0, evals_ = 0;
    // FP64 segments that are their own source and large enough: each pair
    // once (fs_sym.h), one launch per targets-per-lane value
    std::vector<SymSeg> symsegs_;
    int sym_first_[4] = {0, 0, 0, 0}, sym_count_[4] = {0, 0, 0, 0}, sym_items_[4] = {0, 0, 0, 0};
    SymSeg* d_symsegs_ = nullptr;
    size_t o_sympart_ = 0;
    std::vector<SegDev> segs_;
    std::vector<WorkItem> items_;
    std::vector<ScanBlock> scan_;
    BlockPool* pool_ = nullptr;
    DevBuf tables_, work_, items_buf_;
    int grow(DevBuf& b, size_t bytes, hipStream_t stream) {
        return pool_ ? pool_->ensure(b, bytes, stream) : b.ensure(bytes, stream);
    }
    // device views
    SegDev* d_segs_ = nullptr;
    WorkItem* d_items_ = nullptr;
    ScanBlock* d_scan_ = nullptr;
    int32_t* d_seg_of_ = nullptr;
    int32_t* d_tile_seg_ = nullptr;  // mixed: segment of each tile and of each group
    int32_t* d_grp_seg_ = nullptr;
    int32_t* d_unit_seg_ = nullptr;  // mixed: segment of each far-field unit record
    int* d_counter_ = nullptr;
    // work space offsets (bytes into work_)
    size_t o_part_ = 0, o_key_ = 0, o_skey_ = 0, o_perm_ = 0, o_sort_ = 0, o_as_ = 0,
           o_asm1_ = 0, o_ws_ = 0, o_far_ = 0, o_pw_ = 0, o_bsum_ = 0, o_bsa_ = 0, o_tot_ = 0,
           o_gov_ = 0, o_tiles_ = 0, o_groups_ = 0, o_tiles64_ = 0, o_moms_ = 0, o_gexp_ = 0,
           o_gexp64_ = 0,
           o_gmoms_ = 0;
    int ntiles_ = 0, ngroups_ = 0, nrecs_ = 0;
    int lstart_[2] = {0, 0}, nunits_[2] = {0, 0};  // far-field unit records per level
};


}  // namespace fs
