// Host-side discretization: quadrature, mesh, DG matrices, sweep levels.
#include "discretization.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

namespace allium {

void parallel_for(long n, const std::function<void(long, long)> &body)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 2048 || nt == 1) { body(0, n); return; }
    std::vector<std::thread> th;
    const long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        long b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&body, b, e] { body(b, e); });
    }
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------
// Level-symmetric quadrature.  The first-octant ordinates and point weights
// are the published LQn values (the same table the reference's data/lsquad
// files were generated from); point classes are numbered in order of first
// appearance of the sorted ordinate-index triple, directions are ordered
// mu-major then eta within an octant, and octants +++, -++, +-+, --+, ++-,
// -+-, +--, --- (src/angular_quadrature.cpp:114-140 reads them in that order).
// ---------------------------------------------------------------------------
namespace {
struct LQn { int order; std::vector<double> att, wtt; };
const std::vector<LQn> &lqn_table()
{
    static const std::vector<LQn> t = {
        {2, {0.577350269189625764509149}, {1.0}},
        {4, {0.350021174581540677777041, 0.868890300722201205229788}, {0.333333333333333333333333}},
        {6, {0.266635401516704720331535, 0.681507726536546927403750, 0.926180935517489107558380},
         {0.176126130863383433783565, 0.157207202469949899549768}},
        {8, {0.218217890235992381266097, 0.577350269189625764509149, 0.786795792469443145800830,
             0.951189731211341853132399},
         {0.120987654320987654320988, 0.0907407407407407407407407, 0.0925925925925925925925926}},
        {10, {0.189321326478010476671494, 0.508881755582618974382711, 0.694318887594384317279217,
              0.839759962236684758403029, 0.963490981110468484701598},
         {0.0893031479843567214704325, 0.0725291517123655242296233, 0.0450437674364086390490892,
          0.0539281144878369243545650}},
        {12, {0.167212652822713264084504, 0.459547634642594690016761, 0.628019096642130901034766,
              0.760021014833664062877138, 0.872270543025721502340662, 0.971637719251358378302376},
         {0.0707625899700910439766549, 0.0558811015648888075828962, 0.0373376737588285824652402,
          0.0502819010600571181385765, 0.0258512916557503911218290}},
        {14, {0.151985861461031912404799, 0.422156982304796966896263, 0.577350269189625764509149,
              0.698892086775901338963210, 0.802226255231412057244328, 0.893691098874356784901111,
              0.976627152925770351762946},
         {0.0579970408969969964063611, 0.0489007976368104874582568, 0.0227935342411872473257345,
          0.0394132005950078294492985, 0.0380990861440121712365891, 0.0258394076418900119611012,
          0.00826957997262252825269908}},
        {16, {0.138956875067780344591732, 0.392289261444811712294197, 0.537096561300879079878296,
              0.650426450628771770509703, 0.746750573614681064580018, 0.831996556910044145168291,
              0.909285500943725291652116, 0.980500879011739882135849},
         {0.0489872391580385335008367, 0.0413295978698440232405505, 0.0203032007393652080748070,
          0.0265500757813498446015484, 0.0379074407956004002099321, 0.0135295047786756344371600,
          0.0326369372026850701318409, 0.0103769578385399087825920}},
    };
    return t;
}
}  // namespace

bool level_symmetric(int order, int dim, Quadrature &q, std::string &err)
{
    const LQn *set = nullptr;
    for (const auto &s : lqn_table())
        if (s.order == order) set = &s;
    if (!set || dim < 2 || dim > 3) {
        err = "level-symmetric order " + std::to_string(order) + " in " + std::to_string(dim) +
              "D not available (S2..S16, 2D/3D)";
        return false;
    }
    const int M = order / 2;
    // first octant
    std::vector<double> ox, oy, oz, ow;
    std::vector<std::array<int, 3>> classes;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M - i; j++) {
            const int k = M - 1 - i - j;
            std::array<int, 3> key = {i, j, k};
            std::sort(key.begin(), key.end());
            int cls = -1;
            for (size_t c = 0; c < classes.size(); c++)
                if (classes[c] == key) cls = (int)c;
            if (cls < 0) { classes.push_back(key); cls = (int)classes.size() - 1; }
            const double x = set->att[i], y = set->att[j];
            const double z = (M == 1) ? set->att[0] : std::sqrt(1.0 - x * x - y * y);
            ox.push_back(x); oy.push_back(y); oz.push_back(z);
            ow.push_back(set->wtt.at(cls) * (M_PI / 2.0));
        }
    const int noct = dim == 3 ? 8 : 4;
    q.dim = dim;
    q.omega.clear(); q.w.clear();
    for (int o = 0; o < noct; o++) {
        const double sx = (o & 1) ? -1.0 : 1.0, sy = (o & 2) ? -1.0 : 1.0, sz = (o & 4) ? -1.0 : 1.0;
        for (size_t m = 0; m < ox.size(); m++) {
            q.omega.push_back(sx * ox[m]);
            q.omega.push_back(sy * oy[m]);
            q.omega.push_back(dim == 3 ? sz * oz[m] : 0.0);
            q.w.push_back(dim == 3 ? ow[m] : 2.0 * ow[m]);
        }
    }
    q.wsum = 0.0;
    for (double v : q.w) q.wsum += v;
    return true;
}

// ---------------------------------------------------------------------------
// 1D rules and Gauss-Lobatto Lagrange basis on [0,1]
// ---------------------------------------------------------------------------
namespace {
void legendre(int n, double t, double &P, double &dP)
{
    double p0 = 1.0, p1 = t;
    if (n == 0) { P = 1.0; dP = 0.0; return; }
    for (int k = 2; k <= n; k++) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
    }
    P = p1;
    dP = n * (t * p1 - p0) / (t * t - 1.0);
}

struct Rule1D { std::vector<double> x, w; };

Rule1D gauss(int n)
{
    Rule1D r;
    r.x.resize(n); r.w.resize(n);
    for (int i = 0; i < n; i++) {
        double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), P, dP;
        for (int it = 0; it < 100; it++) {
            legendre(n, t, P, dP);
            const double dt = P / dP;
            t -= dt;
            if (std::fabs(dt) < 1e-17) break;
        }
        legendre(n, t, P, dP);
        r.x[n - 1 - i] = 0.5 * (1.0 + t);
        r.w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dP * dP);
    }
    return r;
}

// corner ("lumped") rule: LumpedIntegrationRule, src/lumping.cpp:3-12
Rule1D corners() { return Rule1D{{0.0, 1.0}, {0.5, 0.5}}; }

std::vector<double> gll_nodes(int p)
{
    std::vector<double> x(p + 1);
    if (p == 0) { x[0] = 0.5; return x; }
    x[0] = 0.0; x[p] = 1.0;
    for (int i = 1; i < p; i++) {
        double t = -std::cos(M_PI * i / p), P, dP;
        for (int it = 0; it < 100; it++) {
            legendre(p, t, P, dP);
            const double d2P = (2.0 * t * dP - p * (p + 1.0) * P) / (1.0 - t * t);
            const double dt = dP / d2P;
            t -= dt;
            if (std::fabs(dt) < 1e-17) break;
        }
        x[i] = 0.5 * (1.0 + t);
    }
    return x;
}

struct Basis {
    int dim, p, n;
    std::vector<double> nodes;
    // values and reference gradients of all n tensor functions at xi
    void eval(const double *xi, double *phi, double *dphi) const
    {
        double v[3][8], d[3][8];
        for (int a = 0; a < 3; a++)
            for (int k = 0; k <= p; k++) {
                if (a >= dim) { v[a][k] = (k == 0); d[a][k] = 0.0; continue; }
                double val = 1.0, der = 0.0;
                for (int m = 0; m <= p; m++) {
                    if (m == k) continue;
                    const double r = 1.0 / (nodes[k] - nodes[m]);
                    der = der * (xi[a] - nodes[m]) * r + val * r;
                    val *= (xi[a] - nodes[m]) * r;
                }
                v[a][k] = val; d[a][k] = der;
            }
        const int qz = dim > 2 ? p + 1 : 1, qy = dim > 1 ? p + 1 : 1;
        int i = 0;
        for (int k = 0; k < qz; k++)
            for (int j = 0; j < qy; j++)
                for (int l = 0; l <= p; l++, i++) {
                    phi[i] = v[0][l] * v[1][j] * v[2][k];
                    dphi[3 * i + 0] = d[0][l] * v[1][j] * v[2][k];
                    dphi[3 * i + 1] = v[0][l] * d[1][j] * v[2][k];
                    dphi[3 * i + 2] = v[0][l] * v[1][j] * d[2][k];
                }
    }
};

struct TensorRule { std::vector<std::array<double, 3>> x; std::vector<double> w; };

TensorRule tensor_rule(int dim, const Rule1D &r)
{
    TensorRule t;
    const int n = (int)r.x.size();
    const int nz = dim > 2 ? n : 1, ny = dim > 1 ? n : 1, nx = dim > 0 ? n : 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                t.x.push_back({r.x[i], dim > 1 ? r.x[j] : 0.0, dim > 2 ? r.x[k] : 0.0});
                t.w.push_back(r.w[i] * (dim > 1 ? r.w[j] : 1.0) * (dim > 2 ? r.w[k] : 1.0));
            }
    if (dim == 0) { t.x.push_back({0, 0, 0}); t.w.push_back(1.0); }
    return t;
}

// mfem::IntRules.Get(geom, order) uses order/2 + 1 Gauss points per axis
int npts(int order) { return order / 2 + 1; }

// d-linear map of element e: J[m*3+d] = dX_m/dxi_d
void elem_jacobian(const Mesh &m, int e, const double *xi, double *J)
{
    int c[3];
    m.elem_coords(e, c);
    for (int i = 0; i < 9; i++) J[i] = 0.0;
    const int dim = m.dim;
    for (int cc = 0; cc < (1 << dim); cc++) {
        const int b[3] = {cc & 1, (cc >> 1) & 1, (cc >> 2) & 1};
        const int vi = (c[0] + b[0]) + m.nv[0] * ((dim > 1 ? c[1] + b[1] : 0) + m.nv[1] * (dim > 2 ? c[2] + b[2] : 0));
        const double *xv = &m.vert[(size_t)vi * dim];
        for (int d = 0; d < dim; d++) {
            double v = b[d] ? 1.0 : -1.0;
            for (int d2 = 0; d2 < dim; d2++)
                if (d2 != d) v *= b[d2] ? xi[d2] : 1.0 - xi[d2];
            for (int mm = 0; mm < dim; mm++) J[mm * 3 + d] += v * xv[mm];
        }
    }
}

// det(J) and adj(J) = det * J^{-1}, adj[r*3+m]
double det_adj(int dim, const double *J, double *adj)
{
    if (dim == 2) {
        adj[0] = J[4]; adj[1] = -J[1]; adj[3] = -J[3]; adj[4] = J[0];
        return J[0] * J[4] - J[1] * J[3];
    }
    const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
    adj[0] = e * i - f * h; adj[1] = c * h - b * i; adj[2] = b * f - c * e;
    adj[3] = f * g - d * i; adj[4] = a * i - c * g; adj[5] = c * d - a * f;
    adj[6] = d * h - e * g; adj[7] = b * g - a * h; adj[8] = a * e - b * d;
    return a * adj[0] + b * adj[3] + c * adj[6];
}

// surface measure and outward (un-normalised) normal of element side at xi
double side_ortho(const Mesh &m, int e, int axis, int hi, const double *xi, double *N)
{
    double J[9];
    elem_jacobian(m, e, xi, J);
    const int dim = m.dim;
    int t[2], k = 0;
    for (int d = 0; d < dim; d++)
        if (d != axis) t[k++] = d;
    N[2] = 0.0;
    if (dim == 2) {
        N[0] = J[3 + t[0]];
        N[1] = -J[t[0]];
    } else {
        const double a[3] = {J[t[0]], J[3 + t[0]], J[6 + t[0]]};
        const double b[3] = {J[t[1]], J[3 + t[1]], J[6 + t[1]]};
        N[0] = a[1] * b[2] - a[2] * b[1];
        N[1] = a[2] * b[0] - a[0] * b[2];
        N[2] = a[0] * b[1] - a[1] * b[0];
    }
    double s = 0.0;
    for (int d = 0; d < dim; d++) s += N[d] * J[d * 3 + axis];
    if ((hi && s < 0) || (!hi && s > 0))
        for (int d = 0; d < 3; d++) N[d] = -N[d];
    return std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
}

void side_point(int dim, int axis, int hi, const double *u, double *xi)
{
    int k = 0;
    for (int d = 0; d < 3; d++) xi[d] = d < dim ? (d == axis ? (double)hi : u[k++]) : 0.0;
}
}  // namespace

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
int Mesh::face_of(int e, int side) const
{
    const int ax = side >> 1, hi = side & 1;
    int c[3];
    elem_coords(e, c);
    c[ax] += hi;
    return foff[ax] + c[0] + fext[ax][0] * (c[1] + fext[ax][1] * c[2]);
}

bool build_mesh(int dim, const int *nel, const double *vertices, const double *extents,
                const int *bc, Mesh &m, std::string &err)
{
    if (dim < 2 || dim > 3) { err = "mesh dim must be 2 or 3"; return false; }
    m.dim = dim;
    long nvt = 1, ne = 1;
    for (int d = 0; d < 3; d++) {
        m.nel[d] = d < dim ? nel[d] : 1;
        if (m.nel[d] < 1) { err = "mesh needs >= 1 element per axis"; return false; }
        m.nv[d] = d < dim ? m.nel[d] + 1 : 1;
        nvt *= m.nv[d];
        ne *= m.nel[d];
    }
    if (ne >= (1L << 22)) { err = "mesh exceeds 4,194,303 elements"; return false; }
    m.Ne = (int)ne;
    for (int s = 0; s < 6; s++) m.bc[s] = (s < 2 * dim && bc) ? bc[s] : 0;
    m.vert.resize((size_t)nvt * dim);
    if (vertices) {
        std::memcpy(m.vert.data(), vertices, sizeof(double) * m.vert.size());
    } else {
        for (long v = 0; v < nvt; v++) {
            const long c[3] = {v % m.nv[0], (v / m.nv[0]) % m.nv[1], v / ((long)m.nv[0] * m.nv[1])};
            for (int d = 0; d < dim; d++) m.vert[v * dim + d] = extents[d] * (double)c[d] / m.nel[d];
        }
    }
    for (double x : m.vert)
        if (!std::isfinite(x)) { err = "non-finite vertex coordinate"; return false; }
    // uniform axis-aligned box grid?
    m.uniform = true;
    for (int d = 0; d < dim; d++) {
        const double x0 = m.vert[d];
        const long last = (long)(m.nv[0] - 1) * (d == 0) + (long)m.nv[0] * (m.nv[1] - 1) * (d == 1) +
                          (long)m.nv[0] * m.nv[1] * (m.nv[2] - 1) * (d == 2);
        m.h[d] = (m.vert[last * dim + d] - x0) / m.nel[d];
    }
    for (long v = 0; v < nvt && m.uniform; v++) {
        const long c[3] = {v % m.nv[0], (v / m.nv[0]) % m.nv[1], v / ((long)m.nv[0] * m.nv[1])};
        for (int d = 0; d < dim; d++) {
            const double expect = m.vert[d] + c[d] * m.h[d];
            if (std::fabs(m.vert[v * dim + d] - expect) > 1e-12 * std::fabs(m.h[d])) m.uniform = false;
        }
    }
    // faces
    int f = 0;
    for (int ax = 0; ax < 3; ax++) {
        m.foff[ax] = f;
        for (int d = 0; d < 3; d++) m.fext[ax][d] = m.nel[d] + (d == ax && ax < dim);
        if (ax < dim) f += m.fext[ax][0] * m.fext[ax][1] * m.fext[ax][2];
    }
    m.foff[3] = f;
    m.Nf = f;
    m.f_e1.assign(f, -1); m.f_e2.assign(f, -1); m.f_bside.assign(f, -1);
    m.f_nor.assign((size_t)f * 3, 0.0);
    for (int ax = 0; ax < dim; ax++) {
        const int *ext = m.fext[ax];
        for (int k = 0; k < ext[2]; k++)
            for (int j = 0; j < ext[1]; j++)
                for (int i = 0; i < ext[0]; i++) {
                    const int fid = m.foff[ax] + i + ext[0] * (j + ext[1] * k);
                    int c[3] = {i, j, k};
                    int lo = -1, hi = -1;
                    if (c[ax] > 0) {
                        int cl[3] = {i, j, k};
                        cl[ax]--;
                        lo = cl[0] + m.nel[0] * (cl[1] + m.nel[1] * cl[2]);
                    }
                    if (c[ax] < m.nel[ax]) hi = c[0] + m.nel[0] * (c[1] + m.nel[1] * c[2]);
                    int e1, side;
                    if (lo >= 0 && hi >= 0) { e1 = lo; m.f_e2[fid] = hi; side = 1; }
                    else if (hi >= 0) { e1 = hi; side = 0; m.f_bside[fid] = 2 * ax; }
                    else { e1 = lo; side = 1; m.f_bside[fid] = 2 * ax + 1; }
                    m.f_e1[fid] = e1;
                    const double u[2] = {0.5, 0.5};
                    double xi[3], N[3];
                    side_point(dim, ax, side, u, xi);
                    const double len = side_ortho(m, e1, ax, side, xi, N);
                    if (!(len > 0)) { err = "degenerate face"; return false; }
                    for (int d = 0; d < dim; d++) m.f_nor[(size_t)fid * 3 + d] = N[d] / len;
                }
    }
    return true;
}

// ---------------------------------------------------------------------------
// element and face matrices (AssembleLocalMatrices, src/sweep.cpp:924-1001;
// MGMassIntegrator src/mg_form.cpp:4-65; FaceMassMatricesIntegrator
// src/sweep.cpp:1187-1244; time mass src/sweep.cpp:1003-1022)
// ---------------------------------------------------------------------------
void uniform_element_rule(const Mesh &m, int p, int order, std::vector<double> &phi, std::vector<double> &w)
{
    const int dim = m.dim;
    Basis B;
    B.dim = dim;
    B.p = p;
    B.nodes = gll_nodes(p);
    B.n = 1;
    for (int d = 0; d < dim; d++) B.n *= p + 1;
    const TensorRule r = tensor_rule(dim, gauss(npts(order)));
    double vol = 1.0;
    for (int d = 0; d < dim; d++) vol *= m.h[d];
    phi.assign(r.w.size() * B.n, 0.0);
    w.assign(r.w.size(), 0.0);
    std::vector<double> dphi(3 * B.n);
    for (size_t q = 0; q < r.w.size(); q++) {
        B.eval(r.x[q].data(), &phi[q * B.n], dphi.data());
        w[q] = r.w[q] * vol;
    }
}

bool assemble(const Mesh &m, int p, int lump, ElementData &ed, std::string &err)
{
    if (p < 0 || p > 2) { err = "fe_order must be 0, 1 or 2"; return false; }
    const int dim = m.dim;
    ed.dim = dim; ed.p = p; ed.lump = lump;
    ed.n = 1; ed.nf = 1;
    for (int d = 0; d < dim; d++) ed.n *= p + 1;
    for (int d = 0; d < dim - 1; d++) ed.nf *= p + 1;
    ed.uniform = m.uniform;
    const int n = ed.n, nf = ed.nf;
    Basis B{dim, p, n, gll_nodes(p)};
    // side node lists
    ed.side_nodes.resize(2 * dim * nf);
    for (int s = 0; s < 2 * dim; s++) {
        const int ax = s >> 1, hi = s & 1;
        int k = 0;
        for (int i = 0; i < n; i++) {
            int idx[3] = {i % (p + 1), (i / (p + 1)) % (p + 1), i / ((p + 1) * (p + 1))};
            if (idx[ax] == (hi ? p : 0)) ed.side_nodes[s * nf + k++] = i;
        }
    }
    const TensorRule r_mg = tensor_rule(dim, (lump & 1) ? corners() : gauss(npts(2 * p)));
    const TensorRule r_cv = tensor_rule(dim, (lump & 2) ? corners() : gauss(npts(dim + 2 * p - 1)));
    const TensorRule r_ms = tensor_rule(dim, (lump & 1) ? corners() : gauss(npts(2 * p + dim - 1)));
    const TensorRule r_fc = tensor_rule(dim - 1, (lump & 4) ? corners() : gauss(npts(2 * p)));
    const int nel_asm = m.uniform ? 1 : m.Ne;
    ed.Mu.assign((size_t)nel_asm * n * n, 0.0);
    ed.Mm.assign((size_t)nel_asm * n * n, 0.0);
    ed.Gd.assign((size_t)nel_asm * dim * n * n, 0.0);
    parallel_for(nel_asm, [&](long b, long eend) {
        std::vector<double> phi(n), dphi(3 * n);
        double J[9], adj[9];
        for (long e = b; e < eend; e++) {
            double *Mu = &ed.Mu[(size_t)e * n * n], *Mm = &ed.Mm[(size_t)e * n * n];
            double *Gd = &ed.Gd[(size_t)e * dim * n * n];
            for (size_t q = 0; q < r_mg.w.size(); q++) {
                elem_jacobian(m, (int)e, r_mg.x[q].data(), J);
                const double wq = det_adj(dim, J, adj) * r_mg.w[q];
                B.eval(r_mg.x[q].data(), phi.data(), dphi.data());
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) Mu[i * n + j] += wq * phi[i] * phi[j];
            }
            for (size_t q = 0; q < r_ms.w.size(); q++) {
                elem_jacobian(m, (int)e, r_ms.x[q].data(), J);
                const double wq = det_adj(dim, J, adj) * r_ms.w[q];
                B.eval(r_ms.x[q].data(), phi.data(), dphi.data());
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) Mm[i * n + j] += wq * phi[i] * phi[j];
            }
            for (size_t q = 0; q < r_cv.w.size(); q++) {
                elem_jacobian(m, (int)e, r_cv.x[q].data(), J);
                det_adj(dim, J, adj);
                B.eval(r_cv.x[q].data(), phi.data(), dphi.data());
                for (int d = 0; d < dim; d++)
                    for (int i = 0; i < n; i++) {
                        // det(J) * (grad phi_i)_d = sum_r dphi_{i,r} adj[r][d]
                        double gi = 0.0;
                        for (int rr = 0; rr < dim; rr++) gi += dphi[3 * i + rr] * adj[rr * 3 + d];
                        gi *= -r_cv.w[q];
                        for (int j = 0; j < n; j++) Gd[((size_t)d * n + i) * n + j] += gi * phi[j];
                    }
            }
        }
    });
    // face mass in face-local ordering: F(k,l) = int l_k l_l ds on the side
    auto face_block = [&](int e, int ax, int hi, double *F) {
        std::vector<double> phi(n), dphi(3 * n);
        for (int i = 0; i < nf * nf; i++) F[i] = 0.0;
        const int s = 2 * ax + hi;
        for (size_t q = 0; q < r_fc.w.size(); q++) {
            double xi[3], N[3];
            side_point(dim, ax, hi, r_fc.x[q].data(), xi);
            const double ws = r_fc.w[q] * side_ortho(m, e, ax, hi, xi, N);
            B.eval(xi, phi.data(), dphi.data());
            for (int k = 0; k < nf; k++)
                for (int l = 0; l < nf; l++)
                    F[k * nf + l] += phi[ed.side_nodes[s * nf + k]] * phi[ed.side_nodes[s * nf + l]] * ws;
        }
    };
    if (m.uniform) {
        ed.Fface.assign((size_t)2 * dim * nf * nf, 0.0);
        for (int s = 0; s < 2 * dim; s++) face_block(0, s >> 1, s & 1, &ed.Fface[(size_t)s * nf * nf]);
    } else {
        ed.Fface.assign((size_t)m.Nf * nf * nf, 0.0);
        parallel_for(m.Nf, [&](long b, long fe) {
            for (long f = b; f < fe; f++) {
                int ax = 0;
                while (f >= m.foff[ax + 1]) ax++;
                const int hi = (m.f_e2[f] >= 0) ? 1 : (m.f_bside[f] & 1);
                face_block(m.f_e1[f], ax, hi, &ed.Fface[(size_t)f * nf * nf]);
            }
        });
    }
    for (double v : ed.Mu)
        if (!std::isfinite(v)) { err = "non-finite element matrix"; return false; }
    return true;
}

// ---------------------------------------------------------------------------
// scalar MIP diffusion for DSA (src/moment_discretization.cpp:73-109;
// MIPDiffusionIntegrator::AssembleFaceMatrix, src/moment_integrators.cpp:3-195)
// ---------------------------------------------------------------------------
bool assemble_mip(const Mesh &m, int p, const double *sigma_t, const double *sigma_a, const MipParams &prm,
                  MipMatrix &K, std::string &err)
{
    if (p < 0 || p > 2) { err = "fe_order must be 0, 1 or 2"; return false; }
    const int dim = m.dim;
    int n = 1;
    for (int d = 0; d < dim; d++) n *= p + 1;
    const int ns = 2 * dim, nc = 1 << dim;
    for (int e = 0; e < m.Ne; e++)
        if (!(sigma_t[e] > 0) || !std::isfinite(sigma_a[e])) { err = "MIP needs sigma_t > 0 and finite sigma_a"; return false; }
    K.n = n; K.nside = ns; K.nc = nc;
    K.Kd.assign((size_t)m.Ne * n * n, 0.0);
    K.Ko.assign((size_t)m.Ne * ns * n * n, 0.0);
    K.nb.assign((size_t)m.Ne * ns, -1);
    Basis B{dim, p, n, gll_nodes(p)};
    // vertex -> node interpolation: node coordinates are GLL points on [0,1]
    K.P.assign((size_t)n * nc, 0.0);
    for (int i = 0; i < n; i++) {
        const int idx[3] = {i % (p + 1), (i / (p + 1)) % (p + 1), i / ((p + 1) * (p + 1))};
        for (int c = 0; c < nc; c++) {
            double w = 1.0;
            for (int d = 0; d < dim; d++) {
                const double x = B.nodes[idx[d]];
                w *= ((c >> d) & 1) ? x : 1.0 - x;
            }
            K.P[(size_t)i * nc + c] = w;
        }
    }
    for (int e = 0; e < m.Ne; e++)
        for (int s = 0; s < ns; s++) {
            const int f = m.face_of(e, s);
            const int o = (m.f_e1[f] == e) ? m.f_e2[f] : m.f_e1[f];
            K.nb[(size_t)e * ns + s] = (m.f_e2[f] >= 0) ? o : -1;
        }
    // DiffusionIntegrator and MassIntegrator rules: order 2p + dim - 1
    const TensorRule r_df = tensor_rule(dim, (prm.lump & 2) ? corners() : gauss(npts(2 * p + dim - 1)));
    const TensorRule r_ms = tensor_rule(dim, (prm.lump & 1) ? corners() : gauss(npts(2 * p + dim - 1)));
    // face integrators: order 2p
    const TensorRule r_fc = tensor_rule(dim - 1, (prm.lump & 4) ? corners() : gauss(npts(2 * p)));
    parallel_for(m.Ne, [&](long b, long eend) {
        std::vector<double> phi(n), dphi(3 * n), g(3 * n);
        double J[9], adj[9];
        for (long e = b; e < eend; e++) {
            double *Kd = &K.Kd[(size_t)e * n * n];
            const double D = 1.0 / (3.0 * sigma_t[e]);
            for (size_t q = 0; q < r_df.w.size(); q++) {
                elem_jacobian(m, (int)e, r_df.x[q].data(), J);
                const double det = det_adj(dim, J, adj);
                B.eval(r_df.x[q].data(), phi.data(), dphi.data());
                // det * grad phi_i = dphi_i^T adj
                for (int i = 0; i < n; i++)
                    for (int d = 0; d < dim; d++) {
                        double v = 0.0;
                        for (int r = 0; r < dim; r++) v += dphi[3 * i + r] * adj[r * 3 + d];
                        g[3 * i + d] = v;
                    }
                const double wq = r_df.w[q] * D / det;
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) {
                        double v = 0.0;
                        for (int d = 0; d < dim; d++) v += g[3 * i + d] * g[3 * j + d];
                        Kd[i * n + j] += wq * v;
                    }
            }
            for (size_t q = 0; q < r_ms.w.size(); q++) {
                elem_jacobian(m, (int)e, r_ms.x[q].data(), J);
                const double wq = det_adj(dim, J, adj) * r_ms.w[q] * sigma_a[e];
                B.eval(r_ms.x[q].data(), phi.data(), dphi.data());
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) Kd[i * n + j] += wq * phi[i] * phi[j];
            }
        }
    });
    // faces, one axis at a time so that no two faces of a pass touch the same
    // element block from different threads: faces are split by parity of
    // their position along the axis
    for (int ax = 0; ax < dim; ax++)
        for (int par = 0; par < 2; par++) {
            const int f0 = m.foff[ax], nfa = m.foff[ax + 1] - m.foff[ax];
            parallel_for(nfa, [&](long b, long fend) {
                std::vector<double> phi1(n), phi2(n), dphi1(3 * n), dphi2(3 * n), dn1(n), dn2(n);
                std::vector<double> E((size_t)4 * n * n), Jm((size_t)4 * n * n);
                double J1[9], J2[9], adj1[9], adj2[9], xi1[3], xi2[3], N[3];
                const int n2 = 2 * n;
                for (long fl = b; fl < fend; fl++) {
                    const int f = f0 + (int)fl;
                    const int *ext = m.fext[ax];
                    const int c[3] = {(int)(fl % ext[0]), (int)((fl / ext[0]) % ext[1]), (int)(fl / ((long)ext[0] * ext[1]))};
                    if ((c[ax] & 1) != par) continue;
                    const int e1 = m.f_e1[f], e2 = m.f_e2[f];
                    if (e2 < 0) {
                        // BoundaryMassIntegrator(alpha) on inflow (Marshak) sides
                        const int side = m.f_bside[f];
                        if (m.bc[side] != 0) continue;
                        double *Kd = &K.Kd[(size_t)e1 * n * n];
                        for (size_t q = 0; q < r_fc.w.size(); q++) {
                            side_point(dim, ax, side & 1, r_fc.x[q].data(), xi1);
                            const double ws = r_fc.w[q] * side_ortho(m, e1, ax, side & 1, xi1, N) * prm.alpha;
                            B.eval(xi1, phi1.data(), dphi1.data());
                            for (int i = 0; i < n; i++)
                                for (int j = 0; j < n; j++) Kd[i * n + j] += ws * phi1[i] * phi1[j];
                        }
                        continue;
                    }
                    std::fill(E.begin(), E.end(), 0.0);
                    std::fill(Jm.begin(), Jm.end(), 0.0);
                    const double D1 = 1.0 / (3.0 * sigma_t[e1]), D2 = 1.0 / (3.0 * sigma_t[e2]);
                    for (size_t q = 0; q < r_fc.w.size(); q++) {
                        side_point(dim, ax, 1, r_fc.x[q].data(), xi1);
                        side_point(dim, ax, 0, r_fc.x[q].data(), xi2);
                        const double area = side_ortho(m, e1, ax, 1, xi1, N);   // |N|, N out of e1
                        elem_jacobian(m, e1, xi1, J1);
                        elem_jacobian(m, e2, xi2, J2);
                        const double det1 = det_adj(dim, J1, adj1), det2 = det_adj(dim, J2, adj2);
                        B.eval(xi1, phi1.data(), dphi1.data());
                        B.eval(xi2, phi2.data(), dphi2.data());
                        const double w = r_fc.w[q];
                        const double w1 = w * D1 / (2.0 * det1), w2 = w * D2 / (2.0 * det2);
                        // adj * N (J^{-1} N times det)
                        double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
                        for (int r = 0; r < dim; r++)
                            for (int mm = 0; mm < dim; mm++) { a1[r] += adj1[r * 3 + mm] * N[mm]; a2[r] += adj2[r * 3 + mm] * N[mm]; }
                        for (int j = 0; j < n; j++) {
                            double v1 = 0.0, v2 = 0.0;
                            for (int r = 0; r < dim; r++) { v1 += dphi1[3 * j + r] * a1[r]; v2 += dphi2[3 * j + r] * a2[r]; }
                            dn1[j] = w1 * v1;
                            dn2[j] = w2 * v2;
                        }
                        for (int i = 0; i < n; i++)
                            for (int j = 0; j < n; j++) {
                                E[(size_t)i * n2 + j] += phi1[i] * dn1[j];
                                E[(size_t)i * n2 + n + j] += phi1[i] * dn2[j];
                                E[(size_t)(n + i) * n2 + j] -= phi2[i] * dn1[j];
                                E[(size_t)(n + i) * n2 + n + j] -= phi2[i] * dn2[j];
                            }
                        if (prm.kappa != 0.0) {
                            double wq = (w1 + w2) * area * area * prm.kappa;
                            wq = std::max(wq, prm.lower * w * area);
                            for (int i = 0; i < n; i++)
                                for (int j = 0; j < n; j++) {
                                    Jm[(size_t)i * n2 + j] += wq * phi1[i] * phi1[j];
                                    Jm[(size_t)i * n2 + n + j] -= wq * phi1[i] * phi2[j];
                                    Jm[(size_t)(n + i) * n2 + j] -= wq * phi2[i] * phi1[j];
                                    Jm[(size_t)(n + i) * n2 + n + j] += wq * phi2[i] * phi2[j];
                                }
                        }
                    }
                    // face matrix = sym E^T - E + J, scattered into the element blocks
                    double *K11 = &K.Kd[(size_t)e1 * n * n], *K22 = &K.Kd[(size_t)e2 * n * n];
                    double *K12 = &K.Ko[((size_t)e1 * ns + 2 * ax + 1) * n * n];
                    double *K21 = &K.Ko[((size_t)e2 * ns + 2 * ax) * n * n];
                    for (int i = 0; i < n2; i++)
                        for (int j = 0; j < n2; j++) {
                            const double v = prm.sym * E[(size_t)j * n2 + i] - E[(size_t)i * n2 + j] + Jm[(size_t)i * n2 + j];
                            const int ii = i % n, jj = j % n;
                            if (i < n && j < n) K11[ii * n + jj] += v;
                            else if (i < n) K12[ii * n + jj] += v;
                            else if (j < n) K21[ii * n + jj] += v;
                            else K22[ii * n + jj] += v;
                        }
                }
            });
        }
    for (double v : K.Kd)
        if (!std::isfinite(v)) { err = "non-finite MIP matrix"; return false; }
    return true;
}

bool assemble_p1(const Mesh &m, int p, const double *sigma_t, const double *sigma_a, double alpha, MipMatrix &K,
                 std::string &err, const BlockIpParams *ip)
{
    if (p < 1 || p > 2) { err = "block DSA operators need fe_order 1 or 2"; return false; }
    const int dim = m.dim;
    int n = 1;
    for (int d = 0; d < dim; d++) n *= p + 1;
    const int N = (dim + 1) * n, ns = 2 * dim;
    if (N > 32) { err = "block DSA operators are implemented up to 32 unknowns per element (2D p<=2, 3D p=1)"; return false; }
    if (!(alpha > 0)) { err = "P1SA needs alpha > 0"; return false; }
    for (int e = 0; e < m.Ne; e++)
        if (!(sigma_t[e] > 0) || !std::isfinite(sigma_a[e])) { err = "P1SA needs sigma_t > 0 and finite sigma_a"; return false; }
    K.n = N; K.nside = ns; K.nc = 0;
    K.Kd.assign((size_t)m.Ne * N * N, 0.0);
    K.Ko.assign((size_t)m.Ne * ns * N * N, 0.0);
    K.nb.assign((size_t)m.Ne * ns, -1);
    K.P.clear();
    Basis B{dim, p, n, gll_nodes(p)};
    for (int e = 0; e < m.Ne; e++)
        for (int sd = 0; sd < ns; sd++) {
            const int f = m.face_of(e, sd);
            const int o = (m.f_e1[f] == e) ? m.f_e2[f] : m.f_e1[f];
            K.nb[(size_t)e * ns + sd] = (m.f_e2[f] >= 0) ? o : -1;
        }
    // row (c, i) / column (d, j) of an element block: c n + i; component dim is phi
    auto at = [&](double *blk, int c, int i, int d, int j) -> double & { return blk[(size_t)(c * n + i) * N + d * n + j]; };
    const TensorRule r_v = tensor_rule(dim, gauss(npts(2 * p + dim - 1)));
    const TensorRule r_f = tensor_rule(dim - 1, gauss(npts(2 * p)));
    // VectorMass(3 sigma_t), Mass(sigma_a), D = -(grad psi, J), G = -D^T
    parallel_for(m.Ne, [&](long b, long eend) {
        std::vector<double> phi(n), dphi(3 * n), g(3 * n);
        double J[9], adj[9];
        for (long e = b; e < eend; e++) {
            double *Kd = &K.Kd[(size_t)e * N * N];
            for (size_t q = 0; q < r_v.w.size(); q++) {
                elem_jacobian(m, (int)e, r_v.x[q].data(), J);
                const double det = det_adj(dim, J, adj);
                B.eval(r_v.x[q].data(), phi.data(), dphi.data());
                for (int i = 0; i < n; i++)
                    for (int d = 0; d < dim; d++) {
                        double v = 0.0;
                        for (int r = 0; r < dim; r++) v += dphi[3 * i + r] * adj[r * 3 + d];
                        g[3 * i + d] = v;   // det * physical gradient
                    }
                const double w = r_v.w[q] * std::fabs(det), wg = r_v.w[q] * (det > 0 ? 1.0 : -1.0);
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) {
                        const double mm = w * phi[i] * phi[j];
                        for (int c = 0; c < dim; c++) at(Kd, c, i, c, j) += 3.0 * sigma_t[e] * mm;
                        at(Kd, dim, i, dim, j) += sigma_a[e] * mm;
                        for (int c = 0; c < dim; c++) {
                            const double dv = -wg * g[3 * i + c] * phi[j];   // row psi_i, column J_c,j
                            at(Kd, dim, i, c, j) += dv;
                            at(Kd, c, j, dim, i) -= dv;
                        }
                    }
            }
        }
    });
    // faces, one axis and parity at a time (no two faces of a pass share an element)
    for (int ax = 0; ax < dim; ax++)
        for (int par = 0; par < 2; par++) {
            const int f0 = m.foff[ax], nfa = m.foff[ax + 1] - m.foff[ax];
            parallel_for(nfa, [&](long b, long fend) {
                std::vector<double> ph[2] = {std::vector<double>(n), std::vector<double>(n)}, dph(3 * n);
                double xi[2][3], Nv[3], nh[3];
                for (long fl = b; fl < fend; fl++) {
                    const int f = f0 + (int)fl;
                    const int *ext = m.fext[ax];
                    const int c3[3] = {(int)(fl % ext[0]), (int)((fl / ext[0]) % ext[1]),
                                       (int)(fl / ((long)ext[0] * ext[1]))};
                    if ((c3[ax] & 1) != par) continue;
                    const int e1 = m.f_e1[f], e2 = m.f_e2[f];
                    if (e2 < 0) {
                        const int side = m.f_bside[f], hi = side & 1;
                        const bool marshak = m.bc[side] == 0;
                        if (ip && !marshak) continue;   // the IP blocks have no reflective-side terms
                        const double beta = ip ? 0.0 : (marshak ? 1.0 / (2.0 * alpha) : 1.0 / alpha);
                        double *Kd = &K.Kd[(size_t)e1 * N * N];
                        for (size_t q = 0; q < r_f.w.size(); q++) {
                            side_point(dim, ax, hi, r_f.x[q].data(), xi[0]);
                            const double area = side_ortho(m, e1, ax, hi, xi[0], Nv);
                            for (int d = 0; d < dim; d++) nh[d] = Nv[d] / area;
                            B.eval(xi[0], ph[0].data(), dph.data());
                            const double w = r_f.w[q];
                            for (int i = 0; i < n; i++)
                                for (int j = 0; j < n; j++) {
                                    const double pp = ph[0][i] * ph[0][j];
                                    for (int c = 0; c < dim; c++)
                                        for (int d = 0; d < dim; d++)
                                            at(Kd, c, i, d, j) += beta * w * area * nh[c] * nh[d] * pp;
                                    if (!marshak) continue;
                                    // P1: BoundaryMass(alpha/2) and 0.5 psi J.n; IP: BoundaryMass(alpha) only
                                    at(Kd, dim, i, dim, j) += (ip ? 1.0 : 0.5) * alpha * w * area * pp;
                                    if (ip) continue;
                                    for (int d = 0; d < dim; d++) {
                                        const double dv = 0.5 * w * Nv[d] * pp;   // 0.5 psi J.n
                                        at(Kd, dim, i, d, j) += dv;
                                        at(Kd, d, j, dim, i) -= dv;
                                    }
                                }
                        }
                        continue;
                    }
                    // blocks (a, b) of the face: 0 = e1, 1 = e2
                    double *blk[2][2] = {{&K.Kd[(size_t)e1 * N * N], &K.Ko[((size_t)e1 * ns + 2 * ax + 1) * N * N]},
                                         {&K.Ko[((size_t)e2 * ns + 2 * ax) * N * N], &K.Kd[(size_t)e2 * N * N]}};
                    for (size_t q = 0; q < r_f.w.size(); q++) {
                        side_point(dim, ax, 1, r_f.x[q].data(), xi[0]);
                        side_point(dim, ax, 0, r_f.x[q].data(), xi[1]);
                        const double area = side_ortho(m, e1, ax, 1, xi[0], Nv);   // out of e1
                        for (int d = 0; d < dim; d++) nh[d] = Nv[d] / area;
                        B.eval(xi[0], ph[0].data(), dph.data());
                        B.eval(xi[1], ph[1].data(), dph.data());
                        const double w = r_f.w[q];
                        // interior penalty on phi: P1 alpha/2; IP max(kappa {D} area / det_1, limit)
                        // (PenaltyIntegrator, src/moment_integrators.cpp:206-287)
                        double pen = 0.5 * alpha;
                        if (ip && !ip->ldg) {
                            double J1[9], adj1[9];
                            elem_jacobian(m, e1, xi[0], J1);
                            const double det1 = det_adj(dim, J1, adj1);
                            const double Davg = ip->scale_D ? 0.5 * (1.0 / (3.0 * sigma_t[e1]) + 1.0 / (3.0 * sigma_t[e2])) : 1.0;
                            pen = std::max(Davg * ip->kappa * area / det1, ip->limit);
                        }
                        const double jj = ip ? 0.0 : 0.5 / alpha;   // P1's DGVectorJumpJump
                        for (int a = 0; a < 2; a++)
                            for (int bb = 0; bb < 2; bb++) {
                                const double sab = (a == bb) ? 1.0 : -1.0, sa = a == 0 ? 1.0 : -1.0;
                                double *X = blk[a][bb];
                                // the transpose block of G: rows of element bb, columns of element a
                                double *XT = blk[bb][a];
                                for (int i = 0; i < n; i++)
                                    for (int j = 0; j < n; j++) {
                                        const double pp = ph[a][i] * ph[bb][j];
                                        at(X, dim, i, dim, j) += sab * pen * w * area * pp;
                                        if (jj != 0.0)
                                            for (int c = 0; c < dim; c++)
                                                for (int d = 0; d < dim; d++)
                                                    at(X, c, i, d, j) += sab * jj * w * area * nh[c] * nh[d] * pp;
                                        // [psi]{J.n}; LDG's trace takes element 1's current only
                                        // (LDGTraceIntegrator with beta . N > 0)
                                        const double tw = (ip && ip->ldg) ? (bb == 0 ? 1.0 : 0.0) : 0.5;
                                        if (tw != 0.0)
                                            for (int d = 0; d < dim; d++) {
                                                const double dv = tw * sa * w * Nv[d] * pp;
                                                at(X, dim, i, d, j) += dv;
                                                at(XT, d, j, dim, i) -= dv;
                                            }
                                    }
                            }
                    }
                }
            });
        }
    for (double v : K.Kd)
        if (!std::isfinite(v)) { err = "non-finite block moment matrix"; return false; }
    return true;
}

// ---------------------------------------------------------------------------
// sweep graph levels (src/sweep.cpp:155-208 edges; 308-322, 477-512 Kahn)
// ---------------------------------------------------------------------------
namespace {
int reflect_index(const Quadrature &q, int a, const double *nor, int dim)
{
    double dot = 0.0, Om[3];
    for (int d = 0; d < dim; d++) dot += q.omega[3 * a + d] * nor[d];
    for (int d = 0; d < dim; d++) Om[d] = q.omega[3 * a + d] - 2 * dot * nor[d];
    for (int b = 0; b < q.size(); b++) {
        double diff = 0.0;
        for (int d = 0; d < dim; d++) diff += (q.omega[3 * b + d] - Om[d]) * (q.omega[3 * b + d] - Om[d]);
        if (std::sqrt(diff) < 1e-14) return b;
    }
    return -1;
}
}  // namespace

int build_plan(const Mesh &m, const Quadrature &q, SweepPlan &plan, std::string &err)
{
    const int dim = m.dim, No = q.size(), Ne = m.Ne, ns = 2 * dim;
    if (No >= 1024) { err = "more than 1023 directions"; return -1; }
    plan.level.assign((size_t)Ne * No, -1);
    plan.refl.assign((size_t)No * ns, -1);
    plan.angle_independent = true;
    for (int s = 0; s < ns; s++)
        if (m.bc[s] == 1) plan.angle_independent = false;
    // reflected directions on reflective sides (normals of that side's first face)
    for (int s = 0; s < ns; s++) {
        if (m.bc[s] != 1) continue;
        const int ax = s >> 1;
        for (int f = m.foff[ax]; f < m.foff[ax + 1]; f++) {
            if (m.f_bside[f] != s) continue;
            for (int a = 0; a < No; a++) {
                double dot = 0.0;
                for (int d = 0; d < dim; d++) dot += m.f_nor[(size_t)f * 3 + d] * q.omega[3 * a + d];
                if (dot < 0) {
                    const int ra = reflect_index(q, a, &m.f_nor[(size_t)f * 3], dim);
                    if (ra < 0) { err = "no direction matches a reflection"; return -3; }
                    if (plan.refl[a * ns + s] >= 0 && plan.refl[a * ns + s] != ra) {
                        err = "reflective side is not planar"; return -3;
                    }
                    plan.refl[a * ns + s] = ra;
                }
            }
        }
    }
    // in-degree of each node: interior faces (edge from the upwind side; a zero
    // dot keeps e1 -> e2) and reflective inflow (edge from the reflected angle)
    auto upwind_of = [&](int f, int a) {
        double dot = 0.0;
        for (int d = 0; d < dim; d++) dot += m.f_nor[(size_t)f * 3 + d] * q.omega[3 * a + d];
        return dot;
    };
    std::atomic<long> nedges{0};
    std::atomic<int> cyclic{0};
    if (plan.angle_independent) {
        // Kahn per direction; the frontier index equals the global frontier index
        // of the reference because directions share no edges.
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&] {
                std::vector<int> deg(Ne), cur, nxt;
                long ned = 0;
                for (int a = next++; a < No; a = next++) {
                    std::fill(deg.begin(), deg.end(), 0);
                    for (int f = 0; f < m.Nf; f++)
                        if (m.f_e2[f] >= 0) {
                            ned++;
                            deg[upwind_of(f, a) < 0 ? m.f_e1[f] : m.f_e2[f]]++;
                        }
                    cur.clear();
                    for (int e = 0; e < Ne; e++)
                        if (deg[e] == 0) cur.push_back(e);
                    int L = 0;
                    long done = 0;
                    while (!cur.empty()) {
                        nxt.clear();
                        for (int e : cur) {
                            plan.level[(size_t)e * No + a] = L;
                            done++;
                            for (int s = 0; s < ns; s++) {
                                const int f = m.face_of(e, s);
                                if (m.f_e2[f] < 0) continue;
                                const double dot = upwind_of(f, a);
                                const int up = dot < 0 ? m.f_e2[f] : m.f_e1[f];
                                const int dn = dot < 0 ? m.f_e1[f] : m.f_e2[f];
                                if (up == e && --deg[dn] == 0) nxt.push_back(dn);
                            }
                        }
                        std::swap(cur, nxt);
                        L++;
                    }
                    if (done != Ne) cyclic = 1;
                }
                nedges += ned;
            });
        for (auto &t : th) t.join();
    } else {
        // global Kahn over all (element, angle) nodes
        const long nn = (long)Ne * No;
        std::vector<int> deg(nn, 0);
        long ned = 0;
        for (int f = 0; f < m.Nf; f++) {
            for (int a = 0; a < No; a++) {
                if (m.f_e2[f] >= 0) {
                    deg[(long)(upwind_of(f, a) < 0 ? m.f_e1[f] : m.f_e2[f]) * No + a]++;
                    ned++;
                } else if (m.bc[m.f_bside[f]] == 1 && upwind_of(f, a) < 0) {
                    deg[(long)m.f_e1[f] * No + a]++;
                    ned++;
                }
            }
        }
        nedges = ned;
        std::vector<int> cur, nxt;
        for (long v = 0; v < nn; v++)
            if (deg[v] == 0) cur.push_back((int)v);
        int L = 0;
        long done = 0;
        while (!cur.empty()) {
            nxt.clear();
            for (int v : cur) {
                plan.level[v] = L;
                done++;
                const int e = v / No, a = v % No;
                for (int s = 0; s < ns; s++) {
                    const int f = m.face_of(e, s);
                    if (m.f_e2[f] >= 0) {
                        const double dot = upwind_of(f, a);
                        const int up = dot < 0 ? m.f_e2[f] : m.f_e1[f];
                        const int dn = dot < 0 ? m.f_e1[f] : m.f_e2[f];
                        if (up == e && --deg[(long)dn * No + a] == 0) nxt.push_back(dn * No + a);
                    } else if (m.bc[s] == 1) {
                        // this node is the reflected source of (e, a') for every a'
                        // with refl[a'] == a on this side
                        for (int a2 = 0; a2 < No; a2++)
                            if (plan.refl[a2 * ns + s] == a && upwind_of(f, a2) < 0)
                                if (--deg[(long)e * No + a2] == 0) nxt.push_back(e * No + a2);
                    }
                }
            }
            std::swap(cur, nxt);
            L++;
        }
        if (done != nn) cyclic = 1;
    }
    if (cyclic) { err = "sweep graph is not a DAG"; return -2; }
    plan.nedges = nedges;
    int nlev = 0;
    for (int v : plan.level) nlev = std::max(nlev, v + 1);
    plan.nlevels = nlev;
    // bucket cells by (level, angle, element)
    std::vector<long> cnt((size_t)nlev * No + 1, 0);
    for (int e = 0; e < Ne; e++)
        for (int a = 0; a < No; a++) cnt[(size_t)plan.level[(size_t)e * No + a] * No + a + 1]++;
    for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
    plan.level_ptr.resize(nlev + 1);
    for (int L = 0; L <= nlev; L++) plan.level_ptr[L] = cnt[(size_t)L * No];
    plan.cells.resize((size_t)Ne * No);
    std::vector<long> pos(cnt.begin(), cnt.end() - 1);
    for (int a = 0; a < No; a++)
        for (int e = 0; e < Ne; e++) {
            const size_t key = (size_t)plan.level[(size_t)e * No + a] * No + a;
            plan.cells[pos[key]++] = ((uint32_t)e << 10) | (uint32_t)a;
        }
    return 0;
}

void build_wave_plan(const Mesh &m, const Quadrature &q, SweepPlan &plan, int block, int nparts)
{
    const int Ne = m.Ne, nl = plan.nlevels, No = q.size();
    constexpr int W = 64, NB = 8;
    const int NP = plan.angle_independent ? nparts : 1;
    plan.nparts = NP;
    std::vector<uint8_t> bins(Ne);
    for (int e = 0; e < Ne; e++) {
        int c[3];
        m.elem_coords(e, c);
        int b = 0;
        for (int d = 0; d < 3; d++) b |= ((c[d] / block) & 1) << d;
        bins[e] = (uint8_t)b;
    }
    // part of a direction: in 3D its octant pair (octants differing only in
    // the x sign), in 2D its quadrant
    std::vector<int> part(No, 0);
    if (NP > 1)
        for (int a = 0; a < No; a++) {
            const int oct = (q.omega[3 * a] < 0 ? 1 : 0) | (q.omega[3 * a + 1] < 0 ? 2 : 0) |
                            (q.omega[3 * a + 2] < 0 ? 4 : 0);
            part[a] = (m.dim == 3 ? oct >> 1 : oct) % NP;
        }
    const int NPB = NP * NB;
    auto key_of = [&](int L, int e, int a) {
        return (((size_t)L * NP + part[a]) * NB + bins[e]) * No + a;
    };
    const size_t nkey = (size_t)nl * NPB * No;
    std::vector<long> cnt(nkey, 0);
    for (int e = 0; e < Ne; e++)
        for (int a = 0; a < No; a++) cnt[key_of(plan.level[(size_t)e * No + a], e, a)]++;
    std::vector<long> off(nkey + 1, 0);
    for (size_t k = 0; k < nkey; k++) off[k + 1] = off[k] + (cnt[k] + W - 1) / W * W;
    plan.wcells.assign(off[nkey], SweepPlan::kPad);
    plan.wlevel_base.assign(nl + 1, 0);
    plan.wpb.assign((size_t)nl * (NPB + 1), 0);
    for (int L = 0; L <= nl; L++) plan.wlevel_base[L] = off[(size_t)L * NPB * No];
    for (int L = 0; L < nl; L++) {
        int *w = &plan.wpb[(size_t)L * (NPB + 1)];
        for (int pb = 0; pb < NPB; pb++) {
            const size_t k0 = ((size_t)L * NPB + pb) * No;
            w[pb + 1] = w[pb] + (int)((off[k0 + No] - off[k0]) / W);
        }
    }
    std::vector<long> pos(off.begin(), off.end() - 1);
    for (int a = 0; a < No; a++)
        for (int e = 0; e < Ne; e++)
            plan.wcells[pos[key_of(plan.level[(size_t)e * No + a], e, a)]++] = ((uint32_t)e << 10) | (uint32_t)a;
}

// ---------------------------------------------------------------------------
// FormTransportSource (src/sweep.cpp:1160-1185): DomainLFIntegrator(Q) plus
// BoundaryFlowIntegrator(inflow, Omega, -1, -0.5), both of order 2p
// ---------------------------------------------------------------------------
void form_source(const Mesh &m, const ElementData &ed, const Quadrature &q, int G, const double *Q,
                 const double *inflow, double *source)
{
    const int dim = m.dim, n = ed.n, No = q.size(), Ne = m.Ne, p = ed.p;
    const size_t Ndof = (size_t)Ne * n;
    Basis B{dim, p, n, gll_nodes(p)};
    const TensorRule r_el = tensor_rule(dim, gauss(npts(2 * p)));
    const TensorRule r_fc = tensor_rule(dim - 1, gauss(npts(2 * p)));
    // int phi_i per element (angle independent)
    std::vector<double> ip((size_t)Ne * n, 0.0);
    parallel_for(Ne, [&](long b, long ee) {
        std::vector<double> phi(n), dphi(3 * n);
        double J[9], adj[9];
        for (long e = b; e < ee; e++)
            for (size_t k = 0; k < r_el.w.size(); k++) {
                elem_jacobian(m, (int)e, r_el.x[k].data(), J);
                const double c = r_el.w[k] * det_adj(dim, J, adj);
                B.eval(r_el.x[k].data(), phi.data(), dphi.data());
                for (int i = 0; i < n; i++) ip[(size_t)e * n + i] += c * phi[i];
            }
    });
    for (int g = 0; g < G; g++) {
        for (int a = 0; a < No; a++) {
            double *qa = source + Ndof * ((size_t)a + (size_t)No * g);
            parallel_for(Ne, [&](long b, long ee) {
                for (long e = b; e < ee; e++) {
                    const double Qv = Q ? Q[(size_t)g * Ne + e] : 0.0;
                    for (int i = 0; i < n; i++) qa[(size_t)e * n + i] = Qv * ip[(size_t)e * n + i];
                }
            });
            if (!inflow) continue;
            std::vector<double> phi(n), dphi(3 * n);
            for (int f = 0; f < m.Nf; f++) {
                if (m.f_e2[f] >= 0) continue;
                const double fv = inflow[(size_t)g * 2 * dim + m.f_bside[f]];
                if (fv == 0.0) continue;
                const int e = m.f_e1[f], ax = m.f_bside[f] >> 1, hi = m.f_bside[f] & 1;
                for (size_t k = 0; k < r_fc.w.size(); k++) {
                    double xi[3], N[3];
                    side_point(dim, ax, hi, r_fc.x[k].data(), xi);
                    side_ortho(m, e, ax, hi, xi, N);
                    double un = 0.0;
                    for (int d = 0; d < dim; d++) un += q.omega[3 * a + d] * N[d];
                    if (un >= 0) continue;
                    const double w = -un * r_fc.w[k] * fv;
                    B.eval(xi, phi.data(), dphi.data());
                    for (int i = 0; i < n; i++) qa[(size_t)e * n + i] += w * phi[i];
                }
            }
        }
    }
}

}  // namespace allium
