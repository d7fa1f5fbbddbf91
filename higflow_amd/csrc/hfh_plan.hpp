// Setup-time recorder: walks one projection step in the reference's
// evaluation order (so every point-keyed WLS cache is filled by the same first
// call as in the reference) and turns each point evaluation into either a
// direct state offset (a point that is an unknown's centre) or an affine row
// of an interpolation group.  Linear combinations the step forms right away
// (source average, pressure gradient, 4-point transverse velocity, tensor
// divergence, velocity-gradient tensor, cell divergence, projection gradient)
// are folded into single rows.
#pragma once
#include <map>

#include "hfh_domain.hpp"

namespace hfh {

// reference enums (higflow/src/hig-flow-kernel.h:27-313)
enum { NON_INCREMENTAL = 0, INCREMENTAL = 1 };
enum { NEWTONIAN = 0, VISCOELASTIC = 3 };
enum { EXPLICIT_EULER = 0, SEMI_IMPLICIT_EULER = 3 };
enum { CENTRAL = 0, FIRST_ORDER = 1, SECOND_ORDER = 2 };
enum { MCU = 0, CUBISTA = 1, QUICK = 2 };
enum { OLDROYD_B = 0, GIESEKUS = 1 };
enum { CELL_CENTRAL = 0, CELL_CUBISTA = 1 };

// One interpolation group: rows  out = cst + sum val*X[col]  (columns are
// absolute state offsets of component 0; ncomp components are evaluated
// with in_stride/out_stride).
// identity of a point evaluation: (domain, field offset, point key)
struct DedupKey {
    const void *dom;
    int field;
    int64_t k[3];
    bool operator==(const DedupKey &o) const {
        return dom == o.dom && field == o.field && k[0] == o.k[0] && k[1] == o.k[1] && k[2] == o.k[2];
    }
};
struct DedupHash {
    size_t operator()(const DedupKey &x) const {
        uint64_t h = (uint64_t)(uintptr_t)x.dom * 0x9E3779B97F4A7C15ull ^ (uint64_t)x.field;
        for (int d = 0; d < 3; d++) {
            h ^= (uint64_t)x.k[d] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        }
        return (size_t)h;
    }
};

struct Group {
    std::vector<int> rowptr{0};
    std::vector<int> col;
    std::vector<double> val;
    std::vector<double> rhs0;                            // folded into cst
    std::vector<std::vector<std::pair<int, double>>> bc;  // (bc slot, rhs coefficient)
    std::vector<int> out;  // destination offset (state, or external vector index)
    int ncomp = 1, in_stride = 0;
    std::unordered_map<DedupKey, int, DedupHash> dedup;

    int nrows() const { return (int)out.size(); }
    int add_row(const std::vector<std::pair<int, double>> &cols, double rhs0v,
                const std::vector<std::pair<int, double>> &bcs, int outoff) {
        for (auto &c : cols) {
            col.push_back(c.first);
            val.push_back(c.second);
        }
        rowptr.push_back((int)col.size());
        rhs0.push_back(rhs0v);
        bc.push_back(bcs);
        out.push_back(outoff);
        return (int)out.size() - 1;
    }
    // constants: cst = -(rhs0 + sum coef*bcval)
    std::vector<double> constants(const std::vector<double> &bcval) const {
        std::vector<double> c(out.size());
        for (size_t r = 0; r < out.size(); r++) {
            double rhs = rhs0[r];
            for (auto &b : bc[r]) rhs += b.second * bcval[b.first];
            c[r] = -rhs;
        }
        return c;
    }
};

// Linear combination of point stencils over absolute state offsets.
struct Combo {
    std::vector<std::pair<int, double>> cols;
    std::vector<std::pair<int, double>> bc;
    double rhs0 = 0.0;
    void add(const Stencil &s, int field, double coef) {
        const size_t need = cols.size() + s.ids.size();
        if (cols.capacity() < need) cols.reserve(std::max(need, 2 * cols.capacity()));
        for (size_t k = 0; k < s.ids.size(); k++) add_state(field + s.ids[k], coef * s.w[k]);
        for (auto &b : s.bc) bc.emplace_back(b.first, coef * b.second);
        rhs0 += coef * s.rhs0;
    }
    void add_state(int off, double coef) {
        auto key = [this](size_t i) { return cols[i].first; };
        const int i = index_.find(off, cols.size(), key);
        if (i >= 0) {
            cols[i].second += coef;
            return;
        }
        cols.emplace_back(off, coef);
        index_.appended(off, cols.size(), key);
    }
    void scale(double s) {
        for (auto &e : cols) e.second *= s;
        for (auto &b : bc) b.second *= s;
        rhs0 *= s;
    }

  private:
    IdIndex index_;
};

}  // namespace hfh
