// CPU check of the stencil layout planner (gip_amd/csrc/gip_layout.h):
// reads a CSR operator and an input vector from raw files, plans the
// layout, forms y = A x from the layout exactly as the device kernels do
// (slots in order with clamped columns, then the tail) and from the CSR
// arrays in stored order, and prints one JSON line with the number of rows
// whose results differ bitwise and the layout statistics.  It then forms
// the symmetric form of the slot values (plan_symmetric) as the device
// does: the stored arrays (one per |offset|), the check that every slot of
// negative offset equals its partner row's slot (the device check), and
// the products from the arrays, counted against the full-layout products.
//   layout_check <nrows> <nnz> <rowptr.i32> <col.i32> <val.f64> <x.f64> <ncols_x> [max_tail_frac]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../gip_amd/csrc/gip_layout.h"

template <class T>
static std::vector<T> load(const char *path, size_t n) {
    std::vector<T> v(n);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", n, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: layout_check nrows nnz rowptr col val x ncols_x [max_tail_frac]\n");
        return 2;
    }
    const int32_t nrows = std::atoi(argv[1]), nnz = std::atoi(argv[2]), ncx = std::atoi(argv[7]);
    const double frac = argc > 8 ? std::atof(argv[8]) : 0.05;
    auto rp = load<int32_t>(argv[3], (size_t)nrows + 1);
    auto col = load<int32_t>(argv[4], (size_t)nnz);
    auto val = load<double>(argv[5], (size_t)nnz);
    auto x = load<double>(argv[6], (size_t)ncx);
    gip::StencilPlan P;
    const bool ok = gip::plan_stencil(nrows, nnz, rp.data(), col.data(), P, frac);
    if (!ok) {
        std::printf("{\"planned\": false}\n");
        return 0;
    }
    // every position in range and used once
    std::vector<double> st((size_t)P.stored, 0.0);
    std::vector<char> used((size_t)P.stored, 0);
    long bad_pos = 0;
    for (int32_t j = 0; j < nnz; ++j) {
        const int32_t p = P.pos[(size_t)j];
        if (p < 0 || p >= P.stored || used[(size_t)p]) {
            ++bad_pos;
            continue;
        }
        used[(size_t)p] = 1;
        st[(size_t)p] = val[(size_t)j];
    }
    int32_t cmax = 0;
    for (int32_t c : col) cmax = std::max(cmax, c);
    long bad_rows = 0, bad_cols = 0, one_entry = 0;
    for (const gip::StencilTail &t : P.tail) one_entry += (t.kind & 0xff) == 3;
    for (int32_t r = 0; r < nrows; ++r) {
        double want = 0.0;
        for (int32_t j = rp[(size_t)r]; j < rp[(size_t)r + 1]; ++j) want = want + x[(size_t)col[(size_t)j]] * val[(size_t)j];
        const bool empty_csr = rp[(size_t)r] == rp[(size_t)r + 1];
        const int32_t s = r / 64, lane = r % 64;
        const gip::StencilHdr &h = P.hdr[(size_t)s];
        const gip::StencilTail &tl = P.tail[(size_t)s];
        double got = 0.0;
        bool empty = false;
        for (int32_t k = 0; k < P.w; ++k) {
            int32_t c = r + P.off[(size_t)s * P.w + k];
            c = c < 0 ? 0 : (c > cmax ? cmax : c);
            got = got + x[(size_t)c] * st[(size_t)(((int64_t)s * P.w + k) * 64 + lane)];
        }
        const int kind = tl.kind & 0xff;
        if (kind == 3) {  // the slice's only tail entry: lane, position and column in the record
            const int lt = (tl.kind >> 8) & 63;
            const int32_t t = tl.tsoff, tc = (int32_t)(uint32_t)tl.m0;
            if (__builtin_popcountll(h.m[0]) != 1 || lt != __builtin_ctzll(h.m[0]) || t != h.tsoff ||
                h.m[1] != 0 || h.empty != 0 || t >= (int32_t)P.tcol.size() || tc != P.tcol[(size_t)t])
                ++bad_pos;
            else if (lane == lt) {
                if (tc > cmax)
                    ++bad_cols;
                else
                    got = got + x[(size_t)tc] * st[(size_t)(P.tail_base + t)];
            }
        } else if (kind == 1) {  // first tail entries only: the compact record decides
            const uint64_t bit = 1ull << lane;
            if (tl.m0 & bit) {
                const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
                const int32_t t = tl.tsoff + __builtin_popcountll(tl.m0 & below);
                if (t < 0 || t >= (int32_t)P.tcol.size() || P.tcol[(size_t)t] > cmax)
                    ++bad_cols;
                else
                    got = got + x[(size_t)P.tcol[(size_t)t]] * st[(size_t)(P.tail_base + t)];
            }
        } else if (kind == 2) {
            const uint64_t bit = 1ull << lane;
            empty = (h.empty & bit) != 0;
            for (int j = 0; j < gip::kStencilMaxTail && (h.m[j] & bit); ++j) {
                const int32_t t = gip::stencil_tail_index(h, lane, j);
                if (t < 0 || t >= (int32_t)P.tcol.size() || P.tcol[(size_t)t] > cmax) {
                    ++bad_cols;
                    continue;
                }
                got = got + x[(size_t)P.tcol[(size_t)t]] * st[(size_t)(P.tail_base + t)];
            }
        }
        if ((kind == 0) != (h.any == 0) || tl.tsoff != h.tsoff || (kind != 3 && tl.m0 != h.m[0])) ++bad_pos;
        if (empty != empty_csr || (!empty && std::memcmp(&got, &want, sizeof(double)) != 0)) ++bad_rows;
    }
    // the symmetric form
    gip::SymPlan Y;
    const bool sym_planned = gip::plan_symmetric(P, Y);
    bool sym_ok = false;
    long sym_bad_rows = 0;
    if (sym_planned) {
        const int64_t rows_pad = (int64_t)P.ns * 64;
        std::vector<double> U((size_t)(rows_pad * Y.nu), 0.0);
        auto code = [&](int64_t s, int k) { return (int)Y.code[(size_t)(s * 8 + k)]; };
        for (int64_t r = 0; r < rows_pad; ++r) {
            const int64_t s = r / 64, lane = r % 64;
            for (int j = 0; j < Y.nu; ++j) {
                double v = 0.0;
                for (int k = 0; k < P.w; ++k)
                    if (code(s, k) == j) v = st[(size_t)((s * P.w + k) * 64 + lane)];
                U[(size_t)(j * rows_pad + r)] = v;
            }
        }
        // the value a slot takes in the symmetric product
        auto symval = [&](int32_t r, int k) {
            const int64_t s = r / 64, lane = r % 64;
            const int cd = code(s, k);
            if (cd < 0) return 0.0;
            if (cd < gip::kSymLower) return U[(size_t)(cd * rows_pad + r)];
            const int j = cd - gip::kSymLower;
            const int64_t pr = (int64_t)r - Y.d[j];
            return pr < 0 ? 0.0 : U[(size_t)(j * rows_pad + pr)];
        };
        sym_ok = true;
        for (int32_t r = 0; r < nrows && sym_ok; ++r)
            for (int k = 0; k < P.w; ++k) {
                const double a = st[(size_t)(((int64_t)(r / 64) * P.w + k) * 64 + r % 64)], b = symval(r, k);
                if (std::memcmp(&a, &b, sizeof(double)) != 0) {
                    sym_ok = false;
                    break;
                }
            }
        // slot sums from the symmetric values against the full slot sums
        for (int32_t r = 0; r < nrows; ++r) {
            const int32_t s = r / 64, lane = r % 64;
            double full = 0.0, sym = 0.0;
            for (int32_t k = 0; k < P.w; ++k) {
                int32_t c = r + P.off[(size_t)s * P.w + k];
                c = c < 0 ? 0 : (c > cmax ? cmax : c);
                full = full + x[(size_t)c] * st[(size_t)(((int64_t)s * P.w + k) * 64 + lane)];
                sym = sym + x[(size_t)c] * symval(r, k);
            }
            if (std::memcmp(&full, &sym, sizeof(double)) != 0) ++sym_bad_rows;
        }
    }
    std::printf("{\"planned\": true, \"w\": %d, \"slices\": %d, \"tail\": %zu, \"stored\": %lld, \"bytes\": %lld, "
                "\"bad_pos\": %ld, \"bad_cols\": %ld, \"bad_rows\": %ld, \"sym_planned\": %s, \"sym_nu\": %d, "
                "\"sym_nd\": %d, \"sym_ok\": %s, \"sym_bad_rows\": %ld, \"one_entry_slices\": %ld}\n",
                P.w, P.ns, P.tcol.size(), (long long)P.stored, (long long)gip::stencil_bytes(P, nrows), bad_pos,
                bad_cols, bad_rows, sym_planned ? "true" : "false", Y.nu, Y.nd, sym_ok ? "true" : "false", sym_bad_rows,
                one_entry);
    return 0;
}
