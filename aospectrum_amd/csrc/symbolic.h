// Host symbolic analysis of a localized-orbital bundle: shared CSR pattern of
// H(k)/S(k), fill-reducing ordering on the atom graph and the tile-level
// structure of the static-pivoting LU.  All of it depends on the block table
// only, never on k or on a contour shift, so it runs once per bundle.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace aosx {

struct BlockTable {
    int64_t n_orbitals = 0;
    int64_t n_atoms = 0;
    int64_t n_blocks = 0;
    const int64_t *atom_offsets = nullptr;
    const int32_t *row_atom = nullptr;
    const int32_t *col_atom = nullptr;
    const int32_t *lattice_shift = nullptr;
    const int64_t *block_offsets = nullptr;
};

// Shared CSR pattern with the gather lists of the Bloch sum.  Entry p of the
// CSR is the sum over s in [starts[p], starts[p+1]) of
// value[source[s]] * phase[group[s]]; mirror[p] is the CSR position of the
// transposed entry (p itself on the diagonal).
struct Pattern {
    int64_t n = 0;
    std::vector<int32_t> indptr;   // [n + 1]
    std::vector<int32_t> indices;  // [nnz]
    std::vector<int32_t> rows;     // [nnz] row of every entry
    std::vector<int64_t> source;   // [n_source]
    std::vector<int32_t> group;    // [n_source]
    std::vector<int64_t> starts;   // [nnz + 1]
    std::vector<int32_t> mirror;   // [nnz]
    std::vector<int32_t> shifts;   // [n_groups][3]
    int64_t nnz() const { return (int64_t)indices.size(); }
    int64_t n_groups() const { return (int64_t)shifts.size() / 3; }
};

// Tile-level static structure of L and U (symmetric, since no row exchanges
// are made).  Orbitals are permuted atom by atom into elimination order and
// cut into tiles of T rows; next[k] lists the tiles i > k with L(i,k) != 0,
// which by symmetry are also the tiles j > k with U(k,j) != 0.  When the
// ordering comes in groups (the nodes of a nested dissection), every group
// starts on a tile boundary: the rows between are padding rows with a unit
// diagonal and no couplings, so that no tile mixes two groups.
struct TileStructure {
    int tile = 64;
    int64_t n = 0;
    int64_t n_rows = 0;                   // factor rows: Nt * T, padding included
    int n_diag = 0;                       // Nt
    std::vector<int32_t> atom_order;      // elimination position -> atom
    std::vector<int32_t> perm;            // factor row -> orbital, -1 on padding rows
    std::vector<int32_t> iperm;           // orbital -> factor row
    std::vector<int32_t> padding_rows;    // factor rows that hold no orbital
    std::vector<int32_t> next_ptr;        // [Nt + 1]
    std::vector<int32_t> next_idx;        // concatenated next[k], ascending
    std::vector<int32_t> prev_ptr;        // [Nt + 1]: tiles i < k with k in next[i]
    std::vector<int32_t> prev_idx;
    std::vector<int32_t> tile_id;         // [Nt * Nt], -1 when absent
    int64_t n_tiles = 0;                  // stored tiles (L incl. diagonal + U)
    std::vector<int64_t> entry_offset;    // [nnz] element offset of each CSR entry
    int64_t atom_fill_entries = 0;
    double factor_flops = 0.0;
    double schur_flops = 0.0;
};

// Returns an empty string on success, else the reason (with `code` set).
std::string build_pattern(const BlockTable &blocks, Pattern &out, int &code);
std::string order_atoms(const BlockTable &blocks, int ordering, std::vector<int32_t> &order,
                        int64_t &atom_fill_entries);
// Nested dissection: `order` lists the atoms node by node (leaves and
// separators in elimination order), node g holding order[groups[g] ..
// groups[g+1]).
std::string order_atoms_nd(const BlockTable &blocks, const double *positions, std::vector<int32_t> &order,
                           std::vector<int32_t> &groups);
int64_t atom_fill(const BlockTable &blocks, const std::vector<int32_t> &order);
// `groups` (may be empty) as from order_atoms_nd: each group starts a tile.
std::string build_tiles(const BlockTable &blocks, const Pattern &pattern, const std::vector<int32_t> &atom_order,
                        const std::vector<int32_t> &groups, int tile, TileStructure &out);

}  // namespace aosx
