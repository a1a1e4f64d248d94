// Wannier90 tight-binding input (host code): the _tb.dat and _wsvec.dat
// readers and the real-space operator pipeline of the reference
// (code/src/util/w90util.cpp:40-444, driven by
// code/src/io/TightBindingParameter.cpp:111-148): degeneracy division,
// optional redistribution over the Wigner-Seitz supercell images of the
// wsvec file, symmetrisation H(-R) = H(R)^dagger, optional removal of
// imaginary parts, projection onto the first `dim` lattice directions.
// Quantities are converted to atomic units on reading (eV -> hartree,
// angstrom -> bohr).  Cells are returned in lexicographic order.
#include <array>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/sbewh_w90.h"

namespace {

using cplx = std::complex<double>;
using Cell = std::array<int, 3>;

constexpr double BOHR_A = 5.291772109060855e-1;    // 1 bohr in angstrom
constexpr double HARTREE_EV = 27.211386245771678;  // 1 hartree in eV

thread_local std::string g_err;

struct Ops {
    std::vector<cplx> H, D;  // [n*n], [3][n*n]
    explicit Ops(int n = 0) : H((size_t)n * n), D(3 * (size_t)n * n) {}
};

struct Fail {
    std::string msg;
};

[[noreturn]] void fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Fail{buf};
}

struct File {
    FILE *f;
    explicit File(const char *p) : f(std::fopen(p, "r")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

// Hermitian average of one matrix (W90_mat2D::symmetrize(),
// code/src/util/w90util.cpp:3-15)
void herm_self(cplx *m, int n)
{
    for (int i = 0; i < n; i++) {
        m[i * n + i] = std::real(m[i * n + i]);
        for (int j = i + 1; j < n; j++) {
            const cplx v = 0.5 * (m[i * n + j] + std::conj(m[j * n + i]));
            m[i * n + j] = v;
            m[j * n + i] = std::conj(v);
        }
    }
}

// a(r, c) <- (a(r, c) + conj(b(c, r))) / 2, b(c, r) <- conj of that
// (W90_mat2D::symmetrize(W90_mat2D&), code/src/util/w90util.cpp:17-27)
void herm_pair(cplx *a, cplx *b, int n)
{
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) {
            const cplx v = 0.5 * (a[r * n + c] + std::conj(b[c * n + r]));
            a[r * n + c] = v;
            b[c * n + r] = std::conj(v);
        }
}

void to_real(std::vector<cplx> &m)
{
    for (auto &v : m) v = std::real(v);
}

}  // namespace

struct sbe_w90_tb {
    std::string header;
    double lattice[9] = {};
    int n = 0;
    std::vector<Cell> r;
    std::vector<int> deg;
    std::vector<cplx> H, D;  // [cell][n*n], [cell][3][n*n]
};

struct sbe_w90_wsvec {
    std::string header;
    struct Entry {
        Cell r;
        int m, n;
        std::vector<Cell> T;
    };
    std::vector<Entry> e;
};

struct sbe_w90_model {
    int n = 0;
    double lattice[9] = {};
    std::vector<int> cells;    // [c][3]
    std::vector<double> H, D;  // interleaved complex, [c][n][n], [c][3][n][n]
};

namespace {

template <class F>
int guarded(F &&f)
{
    try {
        f();
        return SBE_W90_OK;
    } catch (const Fail &e) {
        g_err = e.msg;
        return SBE_W90_ERR;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SBE_W90_ERR;
    }
}

// code/src/util/w90util.cpp:40-160
void read_tb(const char *path, sbe_w90_tb &t)
{
    File f(path);
    if (!f.f) fail("Could not open file %s", path);
    char buf[1024];
    if (!std::fgets(buf, sizeof buf, f.f)) fail("Could not read header");
    t.header = buf;
    for (int i = 0; i < 3; i++) {
        double x[3];
        if (std::fscanf(f.f, "%lf %lf %lf", x, x + 1, x + 2) != 3) fail("Could not read lattice vector %d", i);
        for (int k = 0; k < 3; k++) t.lattice[3 * i + k] = x[k] / BOHR_A;
    }
    unsigned nw, nrpts;
    if (std::fscanf(f.f, "%u %u", &nw, &nrpts) != 2) fail("Could not read 'num_wann nrpts'");
    if (nw < 1 || nw > 4096) fail("num_wann %u out of range", nw);
    t.n = (int)nw;
    const size_t nn = (size_t)nw * nw;
    t.r.resize(nrpts);
    t.deg.resize(nrpts);
    t.H.assign(nrpts * nn, 0.0);
    t.D.assign(nrpts * 3 * nn, 0.0);
    for (unsigned i = 0; i < nrpts; i++) {
        unsigned d;
        if (std::fscanf(f.f, "%u", &d) != 1) fail("Could not read degeneracy of Wigner-Seitz point %u", i);
        t.deg[i] = (int)d;
    }
    for (unsigned i = 0; i < nrpts; i++) {
        Cell &r = t.r[i];
        if (std::fscanf(f.f, "%d %d %d", &r[0], &r[1], &r[2]) != 3) fail("Could not read R of cell %u (H)", i);
        for (size_t e = 0; e < nn; e++) {
            int m, n;
            double re, im;
            if (std::fscanf(f.f, "%d %d %lf %lf", &m, &n, &re, &im) != 4)
                fail("Could not read <0m|H|Rn> for R=(%d, %d, %d) at line %zu", r[0], r[1], r[2], e);
            if (m < 1 || m > (int)nw || n < 1 || n > (int)nw) fail("Invalid Wannier indices %d %d", m, n);
            t.H[i * nn + (size_t)(m - 1) * nw + (n - 1)] = cplx(re / HARTREE_EV, im / HARTREE_EV);
        }
    }
    for (unsigned i = 0; i < nrpts; i++) {
        int c[3];
        if (std::fscanf(f.f, "%d %d %d", c, c + 1, c + 2) != 3) fail("Could not read R of cell %u (r)", i);
        for (int k = 0; k < 3; k++)
            if (c[k] != t.r[i][k]) fail("Order of R differs between H and r for Wigner-Seitz point %u", i);
        for (size_t e = 0; e < nn; e++) {
            int m, n;
            double v[6];
            if (std::fscanf(f.f, "%d %d %lf %lf %lf %lf %lf %lf", &m, &n, v, v + 1, v + 2, v + 3, v + 4, v + 5) != 8)
                fail("Could not read <0m|r|Rn> for R=(%d, %d, %d) at line %zu", c[0], c[1], c[2], e);
            if (m < 1 || m > (int)nw || n < 1 || n > (int)nw) fail("Invalid Wannier indices %d %d", m, n);
            for (int d = 0; d < 3; d++)
                t.D[(i * 3 + d) * nn + (size_t)(m - 1) * nw + (n - 1)] =
                    cplx(v[2 * d] / BOHR_A, v[2 * d + 1] / BOHR_A);
        }
    }
}

// code/src/util/w90util.cpp:163-208
void read_wsvec(const char *path, sbe_w90_wsvec &w)
{
    File f(path);
    if (!f.f) fail("Could not open file %s", path);
    char buf[1024];
    if (!std::fgets(buf, sizeof buf, f.f)) fail("Could not read header");
    w.header = buf;
    int r[3];
    unsigned m, n;
    while (std::fscanf(f.f, "%d %d %d %u %u", &r[0], &r[1], &r[2], &m, &n) == 5) {
        sbe_w90_wsvec::Entry e;
        e.r = {r[0], r[1], r[2]};
        if (m < 1 || n < 1) fail("Invalid Wannier indices %u %u", m, n);
        e.m = (int)m - 1;
        e.n = (int)n - 1;
        unsigned cnt;
        if (std::fscanf(f.f, "%u", &cnt) != 1) fail("Could not read supercell offset multiplicity");
        for (unsigned i = 0; i < cnt; i++) {
            if (std::fscanf(f.f, "%d %d %d", &r[0], &r[1], &r[2]) != 3) fail("Could not read superlattice offset vector");
            e.T.push_back({r[0], r[1], r[2]});
        }
        w.e.push_back(std::move(e));
    }
}

Cell add(const Cell &a, const Cell &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Cell neg(const Cell &a) { return {-a[0], -a[1], -a[2]}; }

// code/src/util/w90util.cpp:255-301 (both interpolation schemes)
std::map<Cell, Ops> realspace(const sbe_w90_tb &t, const sbe_w90_wsvec *w)
{
    const int n = t.n;
    const size_t nn = (size_t)n * n;
    std::map<Cell, Ops> res;
    if (!w) {
        for (size_t c = 0; c < t.r.size(); c++) {
            if (t.deg[c] < 1) fail("degeneracy of cell %zu is %d", c, t.deg[c]);
            Ops o(n);
            for (size_t e = 0; e < nn; e++) {
                o.H[e] = t.H[c * nn + e] / (double)t.deg[c];
                for (int d = 0; d < 3; d++) o.D[d * nn + e] = t.D[(c * 3 + d) * nn + e] / (double)t.deg[c];
            }
            res[t.r[c]] = o;
        }
        return res;
    }
    std::map<Cell, std::map<std::pair<int, int>, std::vector<Cell>>> tr;
    for (const auto &e : w->e)
        for (const auto &x : e.T) tr[e.r][{e.m, e.n}].push_back(x);
    for (size_t c = 0; c < t.r.size(); c++) {
        if (t.deg[c] < 1) fail("degeneracy of cell %zu is %d", c, t.deg[c]);
        auto it = tr.find(t.r[c]);
        if (it == tr.end()) fail("wsvec has no entry for R=(%d, %d, %d)", t.r[c][0], t.r[c][1], t.r[c][2]);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                auto jt = it->second.find({i, j});
                if (jt == it->second.end())
                    fail("wsvec has no entry for R=(%d, %d, %d), (%d, %d)", t.r[c][0], t.r[c][1], t.r[c][2], i + 1, j + 1);
                const auto &tv = jt->second;
                for (const Cell &x : tv) {
                    auto ot = res.find(add(t.r[c], x));
                    if (ot == res.end()) ot = res.emplace(add(t.r[c], x), Ops(n)).first;
                    ot->second.H[(size_t)i * n + j] += t.H[c * nn + (size_t)i * n + j] / (double)t.deg[c] / (double)tv.size();
                    for (int d = 0; d < 3; d++)
                        ot->second.D[d * nn + (size_t)i * n + j] +=
                            t.D[(c * 3 + d) * nn + (size_t)i * n + j] / (double)t.deg[c] / (double)tv.size();
                }
            }
    }
    return res;
}

bool negative(const Cell &c) { return c[0] < 0 || (c[0] == 0 && c[1] < 0) || (c[0] == 0 && c[1] == 0 && c[2] < 0); }

// code/src/util/w90util.cpp:303-340
void symmetrize_ops(std::map<Cell, Ops> &rso, int n, bool discard_imag)
{
    const size_t nn = (size_t)n * n;
    for (auto &[ci, r] : rso) {
        if (ci == Cell{0, 0, 0}) {
            herm_self(r.H.data(), n);
            for (int d = 0; d < 3; d++) herm_self(r.D.data() + d * nn, n);
            if (discard_imag) {
                to_real(r.H);
                to_real(r.D);
            }
        }
        if (negative(ci)) {
            auto it = rso.find(neg(ci));
            if (it == rso.end()) fail("no counterpart -R for R=(%d, %d, %d)", ci[0], ci[1], ci[2]);
            Ops &r2 = it->second;
            herm_pair(r.H.data(), r2.H.data(), n);
            for (int d = 0; d < 3; d++) herm_pair(r.D.data() + d * nn, r2.D.data() + d * nn, n);
            if (discard_imag) {
                to_real(r.H);
                to_real(r2.H);
                to_real(r.D);
                to_real(r2.D);
            }
        }
    }
}

// code/src/util/w90util.cpp:417-444: cells beyond dim folded onto their projection
void project(std::map<Cell, Ops> &rso, int dim)
{
    std::map<Cell, Ops> out;
    for (auto &[ci, o] : rso) {
        Cell p = ci;
        for (int d = dim; d < 3; d++) p[d] = 0;
        auto it = out.find(p);
        if (it == out.end()) {
            out.emplace(p, o);
        } else {
            for (size_t e = 0; e < o.H.size(); e++) it->second.H[e] += o.H[e];
            for (size_t e = 0; e < o.D.size(); e++) it->second.D[e] += o.D[e];
        }
    }
    rso.swap(out);
}

}  // namespace

extern "C" {

const char *sbe_w90_last_error(void) { return g_err.c_str(); }

int sbe_w90_tb_read(const char *path, sbe_w90_tb **out)
{
    if (!path || !out) return SBE_W90_ERR;
    *out = nullptr;
    auto *t = new sbe_w90_tb();
    int rc = guarded([&] { read_tb(path, *t); });
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return SBE_W90_OK;
}

void sbe_w90_tb_free(sbe_w90_tb *t) { delete t; }

int sbe_w90_tb_info(const sbe_w90_tb *t, int *num_wann, int *num_cells, double lattice[9])
{
    if (!t) return SBE_W90_ERR;
    if (num_wann) *num_wann = t->n;
    if (num_cells) *num_cells = (int)t->r.size();
    if (lattice) std::memcpy(lattice, t->lattice, sizeof t->lattice);
    return SBE_W90_OK;
}

int sbe_w90_tb_cell(const sbe_w90_tb *t, int c, int r[3], int *degeneracy, double *H, double *D)
{
    if (!t || c < 0 || c >= (int)t->r.size()) return SBE_W90_ERR;
    const size_t nn = (size_t)t->n * t->n;
    if (r)
        for (int k = 0; k < 3; k++) r[k] = t->r[c][k];
    if (degeneracy) *degeneracy = t->deg[c];
    if (H) std::memcpy(H, t->H.data() + c * nn, nn * sizeof(cplx));
    if (D) std::memcpy(D, t->D.data() + c * 3 * nn, 3 * nn * sizeof(cplx));
    return SBE_W90_OK;
}

int sbe_w90_wsvec_read(const char *path, sbe_w90_wsvec **out)
{
    if (!path || !out) return SBE_W90_ERR;
    *out = nullptr;
    auto *w = new sbe_w90_wsvec();
    int rc = guarded([&] { read_wsvec(path, *w); });
    if (rc) {
        delete w;
        return rc;
    }
    *out = w;
    return SBE_W90_OK;
}

void sbe_w90_wsvec_free(sbe_w90_wsvec *w) { delete w; }

int sbe_w90_wsvec_count(const sbe_w90_wsvec *w) { return w ? (int)w->e.size() : SBE_W90_ERR; }

int sbe_w90_wsvec_entry(const sbe_w90_wsvec *w, int i, int r[3], int *m, int *n, int *count, int *T, int capacity)
{
    if (!w || i < 0 || i >= (int)w->e.size()) return SBE_W90_ERR;
    const auto &e = w->e[i];
    if (r)
        for (int k = 0; k < 3; k++) r[k] = e.r[k];
    if (m) *m = e.m;
    if (n) *n = e.n;
    if (count) *count = (int)e.T.size();
    if (T)
        for (int j = 0; j < (int)e.T.size() && j < capacity; j++)
            for (int k = 0; k < 3; k++) T[3 * j + k] = e.T[j][k];
    return SBE_W90_OK;
}

int sbe_w90_model_build(const sbe_w90_tb *tb, const sbe_w90_wsvec *wsvec, int flags, int dim, sbe_w90_model **out)
{
    if (!tb || !out || dim < 0 || dim > 3) return SBE_W90_ERR;
    *out = nullptr;
    auto *m = new sbe_w90_model();
    int rc = guarded([&] {
        const int n = tb->n;
        const size_t nn = (size_t)n * n;
        auto rso = realspace(*tb, wsvec);
        const bool only_real = flags & SBE_W90_ONLY_REAL;
        if (flags & SBE_W90_SYMMETRIZE) {
            symmetrize_ops(rso, n, only_real);
        } else if (only_real) {
            for (auto &[ci, o] : rso) {
                to_real(o.H);
                to_real(o.D);
            }
        }
        if (dim > 0) project(rso, dim);
        if (!rso.count(Cell{0, 0, 0})) fail("the operators do not contain R=(0, 0, 0)");
        m->n = n;
        std::memcpy(m->lattice, tb->lattice, sizeof m->lattice);
        for (auto &[ci, o] : rso) {
            m->cells.insert(m->cells.end(), ci.begin(), ci.end());
            for (size_t e = 0; e < nn; e++) {
                m->H.push_back(o.H[e].real());
                m->H.push_back(o.H[e].imag());
            }
            for (size_t e = 0; e < 3 * nn; e++) {
                m->D.push_back(o.D[e].real());
                m->D.push_back(o.D[e].imag());
            }
        }
    });
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return SBE_W90_OK;
}

int sbe_w90_model_view(const sbe_w90_model *m, sbe_model *view)
{
    if (!m || !view) return SBE_W90_ERR;
    view->num_wann = m->n;
    view->num_cells = (int)(m->cells.size() / 3);
    view->cells = m->cells.data();
    view->H = m->H.data();
    view->D = m->D.data();
    std::memcpy(view->lattice, m->lattice, sizeof view->lattice);
    return SBE_W90_OK;
}

void sbe_w90_model_free(sbe_w90_model *m) { delete m; }

}  // extern "C"
