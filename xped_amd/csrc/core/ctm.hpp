// Corner-transfer-matrix renormalisation (CTMRG) of an iPEPS on a 2x2 unit cell, driven on the GPU.
//
// Follows the reference's CTM<Scalar, Symmetry, TRank = 2, ALL_OUT_LEGS = false> with CORNER
// projection (src/PEPS/CTM.cpp): FROM_A initialisation (:137-240), grow_all (:386-420) with the
// left/right/top/bottom moves (:924-1183), get_projectors (:1185-1337) + decompose
// (PEPSContractions.cpp:16-43), renormalize_* (:1339-1641), contractCorner (:1643-1760),
// computeRDM_h/_v (:528-874) and the nearest-neighbour energy of avg (LinearAlgebra.cpp:198-260,
// CTMSolver.cpp:60-62).  Every tensor operation runs through the block-sparse device layer.
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <vector>

#include "tensor.hpp"

namespace xa {

enum class SvdMode : int {
    Full = 0,        // full FP64 block SVD of every sector, then the reference's truncation rule
    Randomized = 1,  // randomized block SVD with per-sector target ranks (see RankPolicy)
};

// Randomized block SVD parameters: sector q is sketched with
// l_q = min(ceil(rank_factor * k_q) + oversample, m_q, n_q) columns for a target rank k_q.
struct RsvdOptions {
    int oversample = 8;
    double rank_factor = 2.0;
    int power_iters = 1;
    int sketch_precision = XH_FP64;  // GEMM precision of the range-finder products
    double safety = 2.0;             // trust a sketch when safety * sigma_l(B) < sigma_cut
};

struct CTMOptions {
    std::size_t chi = 16;
    int precision = XH_FP64;  // GEMM precision of the contractions
    SvdMode svd = SvdMode::Full;
    RsvdOptions rsvd;  // randomized SVD: panel width, power iterations, sketch precision
    // target rank per sector for randomized SVD; when empty the kept counts of the previous full
    // SVD of the same (site, direction) are used
    std::function<std::map<Q, int>(const DTensor& prod, int site, int dir)> rank_policy;
    bool record_spectra = false;
    // learned ranks: read the network's features from the previous decomposition of the same
    // (site, direction) when it covers every sector (no sketch of its own), else from a fresh sketch
    bool predictor_reuses_spectra = true;
    // run the independent contractions of a move (projector corners, renormalisation chains) in
    // lockstep, each stage's products in one GEMM launch.  Off by default: at D = 8, chi = 256 it
    // halves the contraction GEMM launches (456 -> 264 per step) and their kernel time drops
    // (87 -> 84 ms), but the step takes longer (660 against 625 ms, same process, two runs each)
    bool batch_contractions = false;
    bool check_finite = false;  // check every projector factor for non-finite entries (diagnostics)
    bool verbose = false;       // one line per projector decomposition on stderr
};

struct SpectrumRecord {
    int site = 0, dir = 0;
    std::map<Q, std::vector<double>> sv;      // full spectra per sector (Full mode only)
    std::map<Q, int> kept;                    // kept count per sector
    std::map<Q, std::pair<int, int>> shape;   // block shape per sector
};

// Learned rank prediction for the randomized block SVD.  A cheap sketch (RANK_TOP leading
// singular values per sector) feeds a small network that predicts, for a candidate cut c, how
// many singular values of each sector lie above c; the cut where the predicted counts sum to chi
// (found on two batched grids) gives the target ranks.  Features per sector (rank_features), all decades
// scaled by 1/4:
//   f[0]      log10(s_1 / s_glob)          (s_glob = largest sketched value of the tensor)
//   f[1..15]  log10(s_j / s_1), j = 2..16  (-16 where the sector has fewer values)
//   f[16]     log10(min(m, n)) / 4
//   f[17]     log10(c / s_glob)
//   f[18]     log10(c / s_last)            (s_last = smallest of the sector's sketched values)
// The network output y is log(1 + count).
constexpr int RANK_TOP = 16;
constexpr int RANK_NFEAT = 19;
void rank_features(const double* top, int ntop, int dim, double s_glob, double log_cut, float* f);

struct RandomizedSVD;

struct RankPredictor {
    xh_ctx* ctx = nullptr;
    xh_mlp* net = nullptr;
    int width = 64;
    double margin = 1.25;  // ranks = ceil(margin * predicted count)
    RankPredictor(xh_ctx* c, const float* params, int64_t nparams, int width, double margin);
    ~RankPredictor();
    RankPredictor(const RankPredictor&) = delete;
    RankPredictor& operator=(const RankPredictor&) = delete;
    // target ranks for prod at bond dimension chi (sketch + network + cut search)
    std::map<Q, int> predict(const DTensor& prod, std::size_t chi, std::uint64_t seed) const;
    // predict() of several tensors with their sketches batched into shared grouped launches
    std::vector<std::map<Q, int>> predict_batch(const std::vector<const DTensor*>& prods, std::size_t chi, const std::vector<std::uint64_t>& seeds) const;
    // the network and cut search on given leading singular values per sector (a fresh sketch's,
    // or those of the previous decomposition of the same tensor)
    std::map<Q, int> ranks_from_spectra(const TensorStructure& s, const std::map<Q, std::vector<double>>& spec, std::size_t chi) const;
    // predicted counts above the cut 10^log_cut * s_glob for explicit sketches (row-major nsec x ntop)
    std::vector<double> counts(const std::vector<std::vector<double>>& top, const std::vector<int>& dims, double s_glob, double log_cut) const;
    // the same for several cuts at once (cut-major rows)
    std::vector<double> counts_grid(const std::vector<std::vector<double>>& top, const std::vector<int>& dims, double s_glob,
                                    const std::vector<double>& log_cuts) const;
};

class CTM {
public:
    CTM(Runtime* rt, const std::vector<DTensor>& A, std::size_t chi);
    void set_options(const CTMOptions& o) { opt = o; }
    CTMOptions opt;

    void init_from_A();
    void grow_all();
    void compute_rdm();
    double energy(double Jxy = 1.0, double Jz = 1.0) const;  // (sum E_h + sum E_v) / 4
    // the rank-0 tensors rho.contract(gate) of the 4 horizontal then 4 vertical bonds, whose
    // traces are the bond energies
    std::vector<DTensor> bond_terms(double Jxy = 1.0, double Jz = 1.0) const;
    double step();  // grow_all + computeRDM + energy
    // energy after `steps` more grow_all moves and its gradient with respect to the four site
    // tensors (dA[site], structures of A); the environment on entry is a constant.  Leaves the
    // environment after the moves, as step() would (without the density matrices).
    double energy_gradient(int steps, double Jxy, double Jz, std::vector<DTensor>& dA);

    struct State {
        std::vector<DTensor> C1, C2, C3, C4, T1, T2, T3, T4;
    };
    State state() const;
    void set_state(const State& s);

    const DTensor& A(int x, int y) const { return As[idx(x, y)]; }
    const DTensor& Adag(int x, int y) const { return Adags[idx(x, y)]; }
    std::vector<DTensor> C1, C2, C3, C4, T1, T2, T3, T4, rho_h, rho_v;
    std::vector<SpectrumRecord> spectra;
    std::map<std::pair<int, int>, std::map<Q, int>> last_kept;  // (site, dir) -> kept counts of the last full SVD
    // (site, dir) -> the leading computed singular values per sector of its last decomposition
    std::map<std::pair<int, int>, std::map<Q, std::vector<double>>> last_spectra;
    long n_predicted_from_previous = 0, n_predicted_from_sketch = 0;
    std::shared_ptr<RankPredictor> predictor;  // learned target ranks (randomized mode) when set
    double svd_ms = 0.0;  // host-timed SVD phase (diagnostics)
    long n_rsvd = 0, n_fallback = 0;
    double max_residual_ratio = 0.0;

    static constexpr int Lx = 2, Ly = 2;
    static int idx(int x, int y)
    {
        x = ((x % Lx) + Lx) % Lx;
        y = ((y % Ly) + Ly) % Ly;
        return x * Ly + y;
    }

private:
    Runtime* rt;
    std::vector<DTensor> As, Adags;

    DTensor ct(const DTensor& a, std::vector<int> a1, const DTensor& b, std::vector<int> a2, int rank) const;
    DTensor identity(const std::vector<Qbasis>& dom, const std::vector<Qbasis>& cod) const;
    Qbasis ket(int x, int y, int leg) const;
    Qbasis bra(int x, int y, int leg) const;
    DTensor contract_corner(int x, int y, int corner) const;
    // stage (0..3) of contractCorner for (x, y, corner), given the previous stage's result
    ContractJob corner_stage(int x, int y, int corner, int stage, const DTensor& prev) const;
    // contract_corner of several (x, y, corner) with each stage's products in one launch
    std::vector<DTensor> contract_corners(const std::vector<std::array<int, 3>>& req) const;
    std::vector<std::pair<DTensor, DTensor>> projector_inputs_all(const std::vector<std::pair<int, int>>& xy, int dir);
    std::pair<DTensor, DTensor> decompose(const DTensor& T1, const DTensor& T2, int site, int dir);
    // decompose() of several (T1, T2) pairs of one move; the randomized SVDs run as one batch
    std::vector<std::pair<DTensor, DTensor>> decompose_batch(const std::vector<std::pair<DTensor, DTensor>>& in, const std::vector<int>& sites,
                                                             int dir);
    std::pair<DTensor, DTensor> projector_inputs(int x, int y, int dir);
    std::pair<DTensor, DTensor> get_projectors(int x, int y, int dir);
    // get_projectors for the sites of one move, with their SVDs batched
    void get_projectors_all(const std::vector<std::pair<int, int>>& xy, int dir, std::vector<DTensor>& P1, std::vector<DTensor>& P2);
    void left_move();
    void right_move();
    void top_move();
    void bottom_move();
    DTensor normalized(const DTensor& t) const;
    void keep_spectrum(int site, int dir, const FullSVD& f);
    // a chain of contractions, stage k contracting stage k - 1's result, whose normalized result
    // goes to *out; run_chains advances independent chains in lockstep, each stage's products in
    // one GEMM launch
    struct Chain {
        std::vector<std::function<ContractJob(const DTensor&)>> stages;
        DTensor* out;
    };
    void run_chains(std::vector<Chain>& chains);
    std::vector<DTensor> batched(const std::vector<ContractJob>& jobs) const;  // contract_batch or one by one
    void check_normalized() const;  // throws for a vanished or non-finite norm met by normalized()
};

// Randomized truncated block SVD (Halko, Martinsson & Tropp, SIAM Review 53 (2011), Alg. 4.4,
// subspace iteration with re-orthonormalisation): every sector is sketched with the panel width
// of RsvdOptions for its target rank in `ranks` (0 when absent); then the reference's truncation
// rule over the computed singular values.  `safe` is false when a sector's smallest Ritz value
// is not a safety factor below the smallest kept singular value, i.e. when a singular value the
// sketch missed could belong to the kept set.
struct RandomizedSVD {
    SVDResult res;
    bool safe = true;
    double max_residual_ratio = 0.0;  // max_q sigma_l(B_q) / sigma_cut over the sketched sectors
    std::map<Q, std::vector<double>> ritz;  // every computed singular value per sector (descending)
};
RandomizedSVD tsvd_randomized(const DTensor& t, const std::map<Q, int>& ranks, std::size_t max_keep, double eps_svd, const RsvdOptions& o,
                              std::uint64_t seed);
// The randomized SVDs of several tensors in one set of grouped launches (every GEMM, Cholesky and
// block SVD launch holds the blocks of all of them); result i is what tsvd_randomized(*ts[i],
// ranks[i], ..., seeds[i]) computes, up to the summation order of split-K GEMMs.
std::vector<RandomizedSVD> tsvd_randomized_batch(const std::vector<const DTensor*>& ts, const std::vector<std::map<Q, int>>& ranks,
                                                 std::size_t max_keep, double eps_svd, const RsvdOptions& o, const std::vector<std::uint64_t>& seeds);

}  // namespace xa
