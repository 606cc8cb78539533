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
#pragma once

#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

namespace gip {

constexpr int kStencilMaxW = 8;
constexpr int kStencilSlice = 64;

struct StencilPlan {
    int32_t ns = 0, w = 0;
    std::vector<int32_t> off;     // ns * w offsets
    std::vector<int32_t> tsoff;   // ns + 1: tail start of each slice
    std::vector<int32_t> flag;    // ns: 1 when rows of the slice have a tail or are empty
    std::vector<uint16_t> tinfo;  // nrows: bits 0-7 start in the slice's tail, 8-14 count, 15 empty row
    std::vector<int32_t> tcol;    // tail columns
    std::vector<int32_t> pos;     // per stored entry: position in the value array
    int64_t tail_base = 0;        // ns * w * 64: first tail value
    int64_t stored = 0;           // tail_base + tail entries
};

// Returns false when the operator does not fit the layout: a slice or row
// tail beyond the 8/7-bit fields, more than max_tail_frac of the entries in
// tails, or positions beyond int32.
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
    for (int32_t s = 0; s < ns; ++s)
        for (int32_t k = 0; k < nsel[(size_t)s]; ++k)
            P.off[(size_t)s * w + k] = sel[(size_t)s * kStencilMaxW + k];  // unused slots: offset 0, value 0
    P.tail_base = (int64_t)ns * w * kStencilSlice;
    P.tsoff.assign((size_t)ns + 1, 0);
    P.flag.assign((size_t)ns, 0);
    P.tinfo.assign((size_t)nrows, 0);
    P.tcol.clear();
    P.pos.assign((size_t)nnz, -1);
    const int64_t max_tail = (int64_t)(max_tail_frac * nnz) + kStencilSlice;
    std::vector<int64_t> tail_entry;  // stored-entry index of each tail entry
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t r0 = s * kStencilSlice, r1 = std::min(nrows, r0 + kStencilSlice);
        const int64_t t_s = (int64_t)P.tcol.size();
        const int32_t *off = &P.off[(size_t)s * w];
        for (int32_t r = r0; r < r1; ++r) {
            const int32_t lane = r - r0;
            const int64_t pre = (int64_t)P.tcol.size() - t_s;
            int32_t k = 0, cnt = 0;
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
                P.tcol.push_back(col[j]);
                tail_entry.push_back(j);
                ++cnt;
            }
            const bool empty = rowptr[r] == rowptr[r + 1];
            if (cnt || empty) P.flag[(size_t)s] = 1;
            if ((cnt && pre > 255) || cnt > 127) return false;
            P.tinfo[(size_t)r] = (uint16_t)((cnt ? pre : 0) | (cnt << 8) | (empty ? 0x8000 : 0));
            if ((int64_t)P.tcol.size() > max_tail) return false;
        }
        P.tsoff[(size_t)s + 1] = (int32_t)P.tcol.size();
    }
    P.stored = P.tail_base + (int64_t)P.tcol.size();
    if (P.stored > INT32_MAX) return false;
    for (size_t t = 0; t < tail_entry.size(); ++t) P.pos[(size_t)tail_entry[t]] = (int32_t)(P.tail_base + (int64_t)t);
    return true;
}

// Bytes one FP64 product streams for the operator in this layout (values,
// tail columns, the per-row tail words of flagged slices, slice headers).
inline int64_t stencil_bytes(const StencilPlan &P, int32_t nrows) {
    int64_t flagged_rows = 0;
    for (int32_t s = 0; s < P.ns; ++s)
        if (P.flag[(size_t)s]) flagged_rows += std::min<int64_t>(kStencilSlice, nrows - (int64_t)s * kStencilSlice);
    return 8 * P.stored + 4 * (int64_t)P.tcol.size() + 2 * flagged_rows + (int64_t)P.ns * (4 * P.w + 12);
}

}  // namespace gip
