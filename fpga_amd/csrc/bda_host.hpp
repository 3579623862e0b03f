// Host-side data structures shared by the generator, the analysis (colouring,
// colour-ordered reordering, factor patterns, transposes, tiling) and the
// host code that runs the device kernels.  Plain C++17, no HIP types.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bda {

// Streaming kernels process "tiles": runs of consecutive block rows that
// belong to one colour, hold at most kTileRows rows and at most
// tile_cap(bs, wide) blocks, so that one workgroup can hold the tile's
// values in registers and stage its per-block partial sums in LDS.  Every
// pattern has a narrow tiling (the sweeps) and a wide one with twice the
// blocks per tile (the SpMV).
constexpr int kTileRows = 64;
// A third, tall tiling (narrow block capacity, up to 128 rows) serves sweep
// colours with few blocks per row, whose narrow tiles are short of work.
constexpr int kTileRowsTall = 128;
// The merged launch around the last colour handles at most this many rows
// with blocks in colour C-2 of a second sweep (one per thread group of BS).
constexpr int kTailRows = 64;
// The extended merged launch also handles at most kTailRows rows of colour
// C-3; every row it computes from a packed list carries at most this many
// second-sweep blocks.
constexpr int kTail3Blocks = 4;
inline int tile_cap(int bs, bool wide = false) { return (bs == 3 ? 168 : 128) * (wide ? 2 : 1); }

struct Bcsr {
    int bs = 0;
    int n = 0;
    std::vector<int32_t> rowptr;   // n+1
    std::vector<int32_t> colidx;   // nnzb, sorted per row
    int nnzb() const { return rowptr.empty() ? 0 : rowptr.back(); }
};

// A block pattern stored on the device, with the index of the source block
// each entry copies from (and whether it is transposed).
struct MatPattern {
    int n = 0;
    std::vector<int32_t> rowptr, colidx;
    std::vector<int32_t> src;        // source block index (meaning depends on the matrix)
    std::vector<int32_t> tile_row;   // tile boundaries (row indices), size ntiles+1
    std::vector<int32_t> colour_tile;// first tile of each colour, size ncolours+1
    std::vector<int32_t> tile_row_wide;  // the same for the wide tiling
    std::vector<int32_t> colour_tile_wide;
    std::vector<int32_t> tile_row_tall;  // and for the tall tiling
    std::vector<int32_t> colour_tile_tall;
    int nnzb() const { return rowptr.empty() ? 0 : rowptr.back(); }
    int ntiles() const { return tile_row.empty() ? 0 : (int)tile_row.size() - 1; }
};

struct Analysis {
    int n = 0, bs = 0, ncolours = 0;
    bool level_scheduling = false;
    std::vector<int32_t> colour;      // per original row
    std::vector<int32_t> perm;        // perm[k] = original row at permuted position k
    std::vector<int32_t> iperm;       // iperm[orig] = permuted position
    std::vector<int32_t> colour_ptr;  // ncolours+1
    MatPattern A;    // permuted A; src = original block index
    std::vector<int32_t> diag;        // position of the diagonal block of each permuted row in A
    MatPattern L;    // strictly lower part of permuted A; src = block index in A
    MatPattern U;    // strictly upper part; src = block index in A
    MatPattern At;   // transpose of permuted A; src = block index in A (transposed)
    MatPattern Ut;   // transpose of U (lower); src = block index in U (transposed)
    MatPattern Lt;   // transpose of L (upper); src = block index in L (transposed)
    int max_row_blocks = 0;
};

// First-fit greedy colouring in natural row order on the symmetrised block
// graph (each row takes the smallest colour absent among its already
// coloured neighbours).  level_scheduling selects level sets of the
// lower-triangular graph instead (level = 1 + max level of neighbours j < i).
// Returns the number of colours.
int colour_rows(const Bcsr &m, bool level_scheduling, std::vector<int32_t> &colour);

// Throws std::runtime_error unless m is a block CSR pattern of n > 0 rows:
// rowptr of n + 1 nondecreasing entries from 0 to colidx.size(), and
// strictly increasing column indices in [0, n) within each row.
void validate_pattern(const Bcsr &m);
// Full analysis; throws std::runtime_error on a bad pattern.
void analyse(const Bcsr &m, bool level_scheduling, Analysis &an);
// The same for a matrix whose rows are already in colour order, colour c
// holding rows [colour_ptr[c], colour_ptr[c+1]): no colouring, identity
// permutation; throws when two rows of one colour are coupled.
void analyse_given_order(const Bcsr &m, const std::vector<int32_t> &colour_ptr, Analysis &an);
// The part of the analysis after the colouring (an.colour, an.ncolours set).
void analyse_coloured(const Bcsr &m, Analysis &an);

// Build tiles for pattern p whose rows are grouped by colour_ptr.
void build_tiles(MatPattern &p, const std::vector<int32_t> &colour_ptr, int bs);
void build_tiles(const MatPattern &p, const std::vector<int32_t> &colour_ptr, int cap,
                 std::vector<int32_t> &tile_row, std::vector<int32_t> &colour_tile, int max_rows = kTileRows);

// Synthetic grid --------------------------------------------------------
struct GridSpec {
    int nx = 0, ny = 0, nz = 0, bs = 3;
    uint64_t seed = 0;
    double perm_sigma = 1.0, layer_corr = 0.0, inactive_frac = 0.0;
    double nnc_frac = 0.0;
    int nnc_dist = 3;
    double pert_max = 0.2, upwind = 0.3, diag_noise = 0.05;
};

struct Grid {
    Bcsr m;
    std::vector<double> values;   // nnzb * bs * bs
    std::vector<double> b;        // n * bs
};

void generate_grid(const GridSpec &spec, Grid &g);

// Diagnostic lines on stderr ("INFO: ..." / "ERROR: ...", flushed at
// once), printed when the process-wide debug level is at least `level`.
// The level is read from the environment variable BDA_DEBUG_LEVEL at first
// use and can be changed with bda_set_debug_level; 0 (the default) prints
// nothing.  Level 1: handle creation, analysis and factor summaries, one
// line per solve, errors; level 2 adds per-colour sizes, the launch plan
// and every solve's norm ring.
int debug_level();
void set_debug_level(int level);
void info(int level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace bda
