// Host-side planner of the stencil layout (no device code; also compiled
// into a CPU test that checks the layout against the CSR products).
//
// Rows are grouped in slices of 64 (one wavefront).  For every slice the
// planner picks up to kStencilMaxW column offsets d = col - row that many of
// its rows share (structured meshes: the stencil directions), ascending.
// Slot k of row r then holds the entry of column r + off[k], or 0 when the
// row has no such entry.  Each row's stored entries are matched in stored
// order against the slots, each to a later slot than the previous one; from
// the first entry that cannot be matched on, the row's remaining entries
// form its tail, kept in stored order and summed after the slots.  A product
// that adds the slots in order and then the tail therefore adds the row's
// entries in stored order, with +-0 terms for the slots the row lacks: the
// CSR result, bitwise, whenever the input vector is finite.
//
// Tails are located without per-row words: each slice has a header with the
// start of its tail and one 64-bit lane mask per tail rank j (rows with more
// than j tail entries) plus a mask of its empty rows.  The slice's tail lists
// every row's first tail entry in lane order, then every row's second, and so
// on; lane l's entry j sits at tsoff + (entries of ranks < j) +
// popcount(m[j] & lanes below l).  A wavefront reads the header with scalar
// loads, so a tail entry costs one column load and one gather, as in ELL.
#pragma once

#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

namespace gip {

constexpr int kStencilMaxW = 8;
constexpr int kStencilSlice = 64;
constexpr int kStencilMaxTail = 8;  // tail entries per row

// Per-slice header (80 bytes, read with scalar loads).
struct StencilHdr {
    int32_t tsoff;   // first tail entry of the slice (tail-local index)
    int32_t any;     // 1 when the slice has tail entries or empty rows
    uint64_t m[kStencilMaxTail];  // m[j]: lanes whose row has more than j tail entries
    uint64_t empty;  // lanes whose row has no entries at all
};
static_assert(sizeof(StencilHdr) == 80, "slice header layout");

// Compact per-slice tail record (16 bytes, read for every slice): kind 0 no
// tail and no empty row; kind 1 only first tail entries (lane mask m0, no
// empty row); kind 3 exactly one tail entry in the slice (no empty row):
// its lane in bits 8..13 of `kind`, its position in tsoff and its column in
// m0, so a wavefront has all three from the record; kind 2 anything else
// (the full header decides).  The kind is `kind & 0xff`.
struct StencilTail {
    int32_t tsoff;
    int32_t kind;
    uint64_t m0;
};
static_assert(sizeof(StencilTail) == 16, "slice tail record layout");

struct StencilPlan {
    int32_t ns = 0, w = 0;
    std::vector<int32_t> off;     // ns * w offsets
    std::vector<int8_t> nsel;     // per slice: slots in use (the rest: offset 0, value 0)
    std::vector<StencilHdr> hdr;  // ns slice headers
    std::vector<StencilTail> tail;  // ns compact tail records
    std::vector<int32_t> tcol;    // tail columns
    std::vector<int32_t> pos;     // per stored entry: position in the value array
    int64_t tail_base = 0;        // ns * w * 64: first tail value
    int64_t stored = 0;           // tail_base + tail entries
};

// Returns false when the operator does not fit the layout: a row with more
// than kStencilMaxTail tail entries, more than max_tail_frac of the entries
// in tails, or positions beyond int32.
inline bool plan_stencil(int32_t nrows, int32_t nnz, const int32_t *rowptr, const int32_t *col, StencilPlan &P,
                         double max_tail_frac = 0.05) {
    if (nrows <= 0 || nnz <= 0) return false;
    const int32_t ns = (nrows + kStencilSlice - 1) / kStencilSlice;
    std::vector<int32_t> sel((size_t)ns * kStencilMaxW, 0), nsel((size_t)ns, 0);
    std::vector<int32_t> ds;
    std::vector<std::pair<int32_t, int32_t>> runs;  // (-count, offset)
    int32_t w = 0;
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t r0 = s * kStencilSlice, r1 = std::min(nrows, r0 + kStencilSlice);
        ds.clear();
        for (int32_t r = r0; r < r1; ++r)
            for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) ds.push_back(col[j] - r);
        std::sort(ds.begin(), ds.end());
        runs.clear();
        for (size_t i = 0; i < ds.size();) {
            size_t e = i;
            while (e < ds.size() && ds[e] == ds[i]) ++e;
            runs.push_back({-(int32_t)(e - i), ds[i]});
            i = e;
        }
        // keep offsets shared by at least an eighth of the slice's rows (and two)
        const int32_t need = std::max<int32_t>(2, (r1 - r0) / 8);
        std::sort(runs.begin(), runs.end());
        int32_t k = 0;
        for (const auto &rc : runs) {
            if (k == kStencilMaxW || -rc.first < need) break;
            sel[(size_t)s * kStencilMaxW + k++] = rc.second;
        }
        std::sort(sel.begin() + (size_t)s * kStencilMaxW, sel.begin() + (size_t)s * kStencilMaxW + k);
        nsel[(size_t)s] = k;
        w = std::max(w, k);
    }
    if (w == 0) return false;
    if ((int64_t)ns * w * kStencilSlice > INT32_MAX) return false;
    P.ns = ns;
    P.w = w;
    P.off.assign((size_t)ns * w, 0);
    P.nsel.assign((size_t)ns, 0);
    for (int32_t s = 0; s < ns; ++s) {
        P.nsel[(size_t)s] = (int8_t)nsel[(size_t)s];
        for (int32_t k = 0; k < nsel[(size_t)s]; ++k)
            P.off[(size_t)s * w + k] = sel[(size_t)s * kStencilMaxW + k];  // unused slots: offset 0, value 0
    }
    P.tail_base = (int64_t)ns * w * kStencilSlice;
    P.hdr.assign((size_t)ns, StencilHdr{});
    P.tcol.clear();
    P.pos.assign((size_t)nnz, -1);
    const int64_t max_tail = (int64_t)(max_tail_frac * nnz) + kStencilSlice;
    std::vector<int64_t> tail_entry;  // stored-entry index of each tail entry
    std::vector<int32_t> rt[kStencilSlice];  // this slice's rows: stored-entry indices of their tails
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t r0 = s * kStencilSlice, r1 = std::min(nrows, r0 + kStencilSlice);
        const int32_t *off = &P.off[(size_t)s * w];
        StencilHdr &h = P.hdr[(size_t)s];
        h.tsoff = (int32_t)P.tcol.size();
        for (int32_t r = r0; r < r1; ++r) {
            const int32_t lane = r - r0;
            rt[lane].clear();
            int32_t k = 0;
            bool tail = false;
            for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int32_t d = col[j] - r;
                if (!tail) {
                    int32_t kk = k;
                    while (kk < w && off[kk] != d) ++kk;
                    if (kk < w) {
                        P.pos[(size_t)j] = (int32_t)(((int64_t)s * w + kk) * kStencilSlice + lane);
                        k = kk + 1;
                        continue;
                    }
                    tail = true;
                }
                rt[lane].push_back(j);
            }
            if ((int)rt[lane].size() > kStencilMaxTail) return false;
            for (size_t t = 0; t < rt[lane].size(); ++t) h.m[t] |= 1ull << lane;
            if (rowptr[r] == rowptr[r + 1]) h.empty |= 1ull << lane;
        }
        // the slice's tail: rank 0 entries in lane order, then rank 1, ...
        for (int t = 0; t < kStencilMaxTail; ++t)
            for (int32_t r = r0; r < r1; ++r)
                if ((int)rt[r - r0].size() > t) {
                    const int32_t j = rt[r - r0][(size_t)t];
                    P.tcol.push_back(col[j]);
                    tail_entry.push_back(j);
                }
        h.any = (h.m[0] | h.empty) ? 1 : 0;
        if ((int64_t)P.tcol.size() > max_tail) return false;
    }
    P.tail.assign((size_t)ns, StencilTail{});
    for (int32_t s = 0; s < ns; ++s) {
        const StencilHdr &h = P.hdr[(size_t)s];
        StencilTail &t = P.tail[(size_t)s];
        t.tsoff = h.tsoff;
        t.m0 = h.m[0];
        t.kind = !h.any ? 0 : (h.empty == 0 && h.m[1] == 0) ? 1 : 2;
        if (t.kind == 1 && __builtin_popcountll(h.m[0]) == 1) {
            t.kind = 3 | (__builtin_ctzll(h.m[0]) << 8);
            t.m0 = (uint64_t)(uint32_t)P.tcol[(size_t)h.tsoff];
        }
    }
    P.stored = P.tail_base + (int64_t)P.tcol.size();
    if (P.stored > INT32_MAX) return false;
    for (size_t t = 0; t < tail_entry.size(); ++t) P.pos[(size_t)tail_entry[t]] = (int32_t)(P.tail_base + (int64_t)t);
    return true;
}

// Symmetric form of the slot values.  For an operator whose entries satisfy
// a_{r,r+d} == a_{r+d,r} bitwise, the value of a slot of negative offset -d
// in row r is the value of the slot of offset +d in row r - d, so only the
// slots of offsets >= 0 are stored.  The paired offsets D = {0} and every
// d with -d a slot offset somewhere (ascending, at most kSymMaxD) have one
// array each, indexed by offset: array j holds, for every row r,
// a_{r,r+D[j]}, or 0 when slice r/64 has no slot of that offset.  A
// slice's other slots of offsets >= 0 (a ghost column's, say) take the
// arrays after them, numbered within the slice.  Slot k of slice s reads
// code[s 8 + k]: j < 16 (own row, array j), 16 + j (row r - D[j], array j;
// 0 when r - D[j] < 0) or -1 (an unused slot: 0).  Whether the values
// really are symmetric is checked on the device against the full slot
// values before the symmetric form is used (and again after every change
// of the values); this plan only fixes the structure.
constexpr int kSymMaxD = 4;   // paired offsets
constexpr int kSymMaxU = 8;   // stored arrays
constexpr int kSymLower = 16;  // codes from kSymLower on: a slot of negative offset

struct SymPlan {
    int32_t nd = 0, nu = 0;
    int32_t d[kSymMaxD] = {};
    int64_t own_slots = 0;     // slots of offsets >= 0 over all slices (the values a product streams)
    std::vector<int8_t> code;  // ns * 8
};

inline bool plan_symmetric(const StencilPlan &P, SymPlan &Y) {
    Y = SymPlan{};
    if (P.w <= 0 || P.w > 8 || (int64_t)P.nsel.size() != P.ns) return false;
    std::vector<int32_t> ds{0};
    for (int32_t s = 0; s < P.ns; ++s)
        for (int32_t k = 0; k < P.nsel[(size_t)s]; ++k) {
            const int32_t o = P.off[(size_t)s * P.w + k];
            if (o == INT32_MIN) return false;
            if (o < 0) ds.push_back(-o);
        }
    std::sort(ds.begin(), ds.end());
    ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
    if (ds.size() > (size_t)kSymMaxD) return false;
    Y.nd = (int32_t)ds.size();
    for (int32_t j = 0; j < Y.nd; ++j) Y.d[j] = ds[(size_t)j];
    Y.code.assign((size_t)P.ns * 8, (int8_t)-1);
    int32_t nextra = 0;
    for (int32_t s = 0; s < P.ns; ++s) {
        int32_t e = 0;
        for (int32_t k = 0; k < P.nsel[(size_t)s]; ++k) {
            const int32_t o = P.off[(size_t)s * P.w + k];
            const int32_t a = o < 0 ? -o : o;
            const auto it = std::lower_bound(ds.begin(), ds.end(), a);
            const bool paired = it != ds.end() && *it == a;
            const int32_t j = (int32_t)(it - ds.begin());
            int32_t cd;
            if (o < 0)
                cd = kSymLower + j;
            else if (paired)
                cd = j;
            else
                cd = Y.nd + e++;
            if (o >= 0) ++Y.own_slots;
            Y.code[(size_t)s * 8 + k] = (int8_t)cd;
        }
        nextra = std::max(nextra, e);
    }
    Y.nu = Y.nd + nextra;
    return Y.nu <= kSymMaxU;
}

// Position of tail entry `rank` of lane `lane` in the slice's tail
// (tail-local index), from the slice header.
inline int32_t stencil_tail_index(const StencilHdr &h, int lane, int rank) {
    int32_t t = h.tsoff;
    for (int j = 0; j < rank; ++j) t += __builtin_popcountll(h.m[j]);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    return t + __builtin_popcountll(h.m[rank] & below);
}

// Bytes one FP64 product streams for the operator in this layout (values,
// tail columns, slice offsets, tail records, the full headers of kind-2
// slices).
inline int64_t stencil_bytes(const StencilPlan &P, int32_t) {
    int64_t general = 0;
    for (const StencilTail &t : P.tail) general += (t.kind & 0xff) == 2;
    return 8 * P.stored + 4 * (int64_t)P.tcol.size() + (int64_t)P.ns * (4 * P.w + (int64_t)sizeof(StencilTail)) +
           general * (int64_t)sizeof(StencilHdr);
}

}  // namespace gip
