// scene.cpp -- host-side scene construction and frame output of libmpm_hip.
//
// * Object sampling after MPMSimulator.cpp:198-276: model matrix M = T*S*R
//   (model.cpp:202-206, glm translate -> scale -> rotate), world AABB of the
//   transformed shape with the grid-range checks of MPMSimulator.cpp:210-222,
//   `ppc` jittered samples per cell of the AABB (:239-249), each kept when
//   M^-1 p lies inside the shape (:249-251).
// * Point-in-mesh: parity of +z ray crossings (the method of mesh_query's
//   point_inside_mesh, thirdparties/mesh_query0.1/mesh_query.cpp:303), with
//   triangles bucketed on a 2D grid over x-y.  Orientations are exact (sign
//   of a floating-point expansion when the plain double value is too close to
//   zero to trust), and a ray through an edge or a vertex of the projected
//   mesh is decided by the half-open rules of mesh_query's tri_zcast
//   (mesh_query.cpp:18-111), so every crossing is counted once.
// * OBJ reader (v / f records only) standing in for tinyobj.
// * BGEO writer standing in for Partio::write (FilesIO.cpp:6-23).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mpm/mpm.h"

namespace mpm {

int set_error(int code, const std::string &msg);

// ---------------------------------------------------------------------------
// Counter-based jitter: SplitMix64 finaliser of (seed, sample index, axis).
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline float unit_float(uint64_t seed, uint64_t idx, uint32_t axis)
{
    uint64_t h = mix64(seed ^ mix64(idx * 4 + axis));
    return (float)(h >> 40) * (1.0f / 16777216.0f);    // [0, 1)
}

// ---------------------------------------------------------------------------
// 4x4 affine transform in double, column vectors: world = M * local.
struct Affine {
    double m[3][4];
};

static Affine model_matrix(const float t[3], const float s[3], const float axis[3], float deg)
{
    double ax = axis[0], ay = axis[1], az = axis[2];
    double n = sqrt(ax * ax + ay * ay + az * az);
    if (n > 0) { ax /= n; ay /= n; az /= n; } else { ax = 1; ay = 0; az = 0; }
    double th = (double)deg * M_PI / 180.0, c = cos(th), sn = sin(th), oc = 1.0 - c;
    double R[3][3] = {{c + oc * ax * ax, oc * ax * ay - sn * az, oc * ax * az + sn * ay},
                      {oc * ay * ax + sn * az, c + oc * ay * ay, oc * ay * az - sn * ax},
                      {oc * az * ax - sn * ay, oc * az * ay + sn * ax, c + oc * az * az}};
    Affine M;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) M.m[r][k] = s[r] * R[r][k];   // S * R
        M.m[r][3] = t[r];                                         // T * (S * R)
    }
    return M;
}

static void apply(const Affine &M, const double p[3], double out[3])
{
    for (int r = 0; r < 3; ++r) out[r] = M.m[r][0] * p[0] + M.m[r][1] * p[1] + M.m[r][2] * p[2] + M.m[r][3];
}

static bool invert(const Affine &M, Affine &I)
{
    const double(*a)[4] = M.m;
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0.0 || !isfinite(det)) return false;
    double id = 1.0 / det;
    double B[3][3];
    B[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
    B[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    B[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    B[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
    B[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    B[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    B[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
    B[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    B[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) I.m[r][k] = B[r][k];
        I.m[r][3] = -(B[r][0] * a[0][3] + B[r][1] * a[1][3] + B[r][2] * a[2][3]);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Exact orientation signs.  A sum of products of doubles is accumulated
// without rounding as a floating-point expansion (Shewchuk 1997, "Adaptive
// precision floating-point arithmetic and fast robust geometric predicates":
// Knuth's two-sum, the exact product a*b = p + fma(a, b, -p), and
// grow-expansion with zero elimination); the components stay nonoverlapping
// and ordered by magnitude, so the sign of the sum is the sign of the last.
struct Expansion {
    double c[128];
    int n = 0;
    void add(double b)
    {
        double q = b;
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const double x = q + c[i];
            const double bv = x - q, av = x - bv;
            const double err = (q - av) + (c[i] - bv);
            q = x;
            if (err != 0.0) c[m++] = err;
        }
        if (q != 0.0) c[m++] = q;
        n = m;
    }
    void add_product(double a, double b)              // + a b
    {
        const double p = a * b;
        add(std::fma(a, b, -p));
        add(p);
    }
    void add_product3(double a, double b, double d)   // + a b d
    {
        const double p = a * b, pe = std::fma(a, b, -p);
        const double q = p * d, qe = std::fma(p, d, -q);
        const double r = pe * d, re = std::fma(pe, d, -r);
        add(qe); add(re); add(r); add(q);
    }
    int sign() const { return n == 0 ? 0 : (c[n - 1] > 0.0 ? 1 : -1); }
};

// sign of (b - a) x (c - a) in the x-y plane: +1 when a, b, c turn
// counter-clockwise
static int orient2d(const double *a, const double *b, const double *c)
{
    const double l = (b[0] - a[0]) * (c[1] - a[1]), r = (b[1] - a[1]) * (c[0] - a[0]);
    const double det = l - r;
    // far above the rounding error of the double evaluation (< 4e-16 of the
    // permanent): its sign is exact
    if (fabs(det) > 1e-12 * (fabs(l) + fabs(r))) return det > 0.0 ? 1 : -1;
    Expansion e;
    e.add_product(b[0], c[1]); e.add_product(-b[0], a[1]); e.add_product(-a[0], c[1]);
    e.add_product(-b[1], c[0]); e.add_product(b[1], a[0]); e.add_product(a[1], c[0]);
    return e.sign();
}

static double det3(const double *u, const double *v, const double *w)
{
    return u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
}

static void add_det3(Expansion &e, double sg, const double *u, const double *v, const double *w)
{
    e.add_product3(sg * u[0], v[1], w[2]); e.add_product3(-sg * u[0], v[2], w[1]);
    e.add_product3(-sg * u[1], v[0], w[2]); e.add_product3(sg * u[1], v[2], w[0]);
    e.add_product3(sg * u[2], v[0], w[1]); e.add_product3(-sg * u[2], v[1], w[0]);
}

// sign of det[a - d; b - d; c - d] (rows), i.e. of (a - d) . ((b - d) x (c - d))
static int orient3d(const double *a, const double *b, const double *c, const double *d)
{
    double A[3], B[3], C[3];
    for (int k = 0; k < 3; ++k) { A[k] = a[k] - d[k]; B[k] = b[k] - d[k]; C[k] = c[k] - d[k]; }
    const double det = det3(A, B, C);
    const double perm = fabs(A[0]) * (fabs(B[1] * C[2]) + fabs(B[2] * C[1])) +
                        fabs(A[1]) * (fabs(B[0] * C[2]) + fabs(B[2] * C[0])) +
                        fabs(A[2]) * (fabs(B[0] * C[1]) + fabs(B[1] * C[0]));
    if (fabs(det) > 1e-12 * perm) return det > 0.0 ? 1 : -1;
    // det of [[a, 1], [b, 1], [c, 1], [d, 1]] = |abc| - |abd| + |acd| - |bcd|
    Expansion e;
    add_det3(e, 1.0, a, b, c);
    add_det3(e, -1.0, a, b, d);
    add_det3(e, 1.0, a, c, d);
    add_det3(e, -1.0, b, c, d);
    return e.sign();
}

// Does the ray p + t (0, 0, 1), t > 0, cross triangle (q, r, s)?  A triangle
// whose x-y projection is degenerate is never crossed; one the ray touches
// only at or below p is not crossed.  A ray through the projection's edge or
// vertex is counted by half-open rules on the projected, counter-clockwise
// triangle (those of mesh_query's tri_zcast, mesh_query.cpp:18-111): on edge
// u -> v it crosses when u lies above v (or level with v and to its left);
// at vertex w (incoming from u, outgoing to v) when w.y >= v.y and
// w.y < u.y.  Two triangles sharing an edge see it in opposite directions
// when they lie on both sides of it in projection, so exactly one counts it.
static bool ray_crosses(const double *p, const double *q, const double *r, const double *s)
{
    const int o = orient2d(q, r, s);
    if (o == 0) return false;
    if (o < 0) std::swap(r, s);
    if (orient3d(p, q, r, s) >= 0) return false;       // p on or above the triangle's plane
    const int a = orient2d(p, q, r), b = orient2d(p, r, s), c = orient2d(p, s, q);
    if (a < 0 || b < 0 || c < 0) return false;
    const int zeros = (a == 0) + (b == 0) + (c == 0);
    if (zeros == 0) return true;
    auto edge = [](const double *u, const double *v) { return u[1] > v[1] || (u[1] == v[1] && u[0] < v[0]); };
    if (zeros == 1) return a == 0 ? edge(q, r) : (b == 0 ? edge(r, s) : edge(s, q));
    // on a vertex: the one shared by the two zero edges
    if (a == 0 && c == 0) return q[1] >= r[1] && q[1] < s[1];
    if (a == 0 && b == 0) return r[1] >= s[1] && r[1] < q[1];
    return s[1] >= q[1] && s[1] < r[1];
}

// ---------------------------------------------------------------------------
// Triangle mesh with a 2D (x-y) bucket grid for +z ray parity queries.
struct Mesh {
    std::vector<double> v;      // 3 per vertex
    std::vector<int> f;         // 3 per triangle
    double lo[3], hi[3];
    int gx = 1, gy = 1;
    std::vector<std::vector<int>> cells;

    void build()
    {
        for (int a = 0; a < 3; ++a) { lo[a] = 1e300; hi[a] = -1e300; }
        for (size_t i = 0; i < v.size(); i += 3)
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[i + a]); hi[a] = std::max(hi[a], v[i + a]); }
        int nt = (int)(f.size() / 3);
        int g = std::max(1, std::min(256, (int)sqrt((double)nt)));
        gx = gy = g;
        cells.assign((size_t)gx * gy, {});
        for (int t = 0; t < nt; ++t) {
            double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
            for (int k = 0; k < 3; ++k) {
                const double *p = &v[3 * f[3 * t + k]];
                x0 = std::min(x0, p[0]); x1 = std::max(x1, p[0]);
                y0 = std::min(y0, p[1]); y1 = std::max(y1, p[1]);
            }
            // cell_x / cell_y are monotone, so a triangle whose closed
            // projection holds p is listed in p's cell
            int cx0 = cell_x(x0), cx1 = cell_x(x1), cy0 = cell_y(y0), cy1 = cell_y(y1);
            for (int cy = cy0; cy <= cy1; ++cy)
                for (int cx = cx0; cx <= cx1; ++cx) cells[(size_t)cy * gx + cx].push_back(t);
        }
    }
    int cell_x(double x) const
    {
        double w = hi[0] - lo[0];
        int c = w > 0 ? (int)((x - lo[0]) / w * gx) : 0;
        return std::max(0, std::min(gx - 1, c));
    }
    int cell_y(double y) const
    {
        double w = hi[1] - lo[1];
        int c = w > 0 ? (int)((y - lo[1]) / w * gy) : 0;
        return std::max(0, std::min(gy - 1, c));
    }
    // Odd number of triangles crossed by the ray p + t (0,0,1), t > 0.
    bool inside(const double p[3]) const
    {
        if (p[0] < lo[0] || p[0] > hi[0] || p[1] < lo[1] || p[1] > hi[1] || p[2] < lo[2] || p[2] > hi[2])
            return false;
        const std::vector<int> &c = cells[(size_t)cell_y(p[1]) * gx + cell_x(p[0])];
        int crossings = 0;
        for (int t : c)
            crossings += ray_crosses(p, &v[3 * f[3 * t]], &v[3 * f[3 * t + 1]], &v[3 * f[3 * t + 2]]) ? 1 : 0;
        return (crossings & 1) != 0;
    }
};

static int parse_index(const std::string &tok, int nverts)
{
    int k = atoi(tok.c_str());   // "a", "a/b", "a//c", "a/b/c": leading integer
    if (k < 0) k = nverts + k + 1;
    return k - 1;
}

static int load_obj(const char *path, Mesh &mesh)
{
    std::ifstream in(path);
    if (!in) return set_error(MPM_ERR_IO, std::string("cannot open OBJ file: ") + (path ? path : "(null)"));
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            double x, y, z;
            ls >> x >> y >> z;
            mesh.v.push_back(x); mesh.v.push_back(y); mesh.v.push_back(z);
        } else if (tag == "f") {
            std::vector<int> idx;
            std::string tok;
            int nv = (int)(mesh.v.size() / 3);
            while (ls >> tok) idx.push_back(parse_index(tok, nv));
            for (size_t k = 2; k < idx.size(); ++k) {   // fan triangulation
                mesh.f.push_back(idx[0]); mesh.f.push_back(idx[k - 1]); mesh.f.push_back(idx[k]);
            }
        }
    }
    int nv = (int)(mesh.v.size() / 3);
    if (nv == 0 || mesh.f.empty()) return set_error(MPM_ERR_IO, std::string("OBJ file has no triangles: ") + path);
    for (int i : mesh.f)
        if (i < 0 || i >= nv) return set_error(MPM_ERR_IO, std::string("OBJ face index out of range: ") + path);
    mesh.build();
    return MPM_OK;
}

// Inside test of the sampler for n points given in the OBJ file's own
// coordinates (the frame in which sample_object tests M^-1 p).
int points_in_mesh(const char *path, const double *pts, int64_t n, uint8_t *inside)
{
    Mesh mesh;
    int rc = load_obj(path, mesh);
    if (rc) return rc;
    for (int64_t i = 0; i < n; ++i) inside[i] = mesh.inside(pts + 3 * i) ? 1 : 0;
    return MPM_OK;
}

// ---------------------------------------------------------------------------
int sample_object(const mpm_config &cfg, const mpm_object &obj, std::vector<float> &xyz)
{
    xyz.clear();
    Affine M = model_matrix(obj.translation, obj.scale, obj.rotation_axis, obj.rotation_deg), Minv;
    if (!invert(M, Minv)) return set_error(MPM_ERR_INVALID, "object transform is singular");
    Mesh mesh;
    std::vector<double> corners;   // local points whose images bound the object
    if (obj.shape == MPM_SHAPE_MESH) {
        int rc = load_obj(obj.obj_path, mesh);
        if (rc) return rc;
        corners = mesh.v;
    } else if (obj.shape == MPM_SHAPE_BOX || obj.shape == MPM_SHAPE_SPHERE) {
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            if (obj.shape == MPM_SHAPE_BOX) { lo[a] = obj.shape_param[a]; hi[a] = obj.shape_param[3 + a]; }
            else { lo[a] = obj.shape_param[a] - obj.shape_param[3]; hi[a] = obj.shape_param[a] + obj.shape_param[3]; }
            if (!(hi[a] > lo[a])) return set_error(MPM_ERR_INVALID, "empty shape");
        }
        for (int k = 0; k < 8; ++k) {
            corners.push_back((k & 1) ? hi[0] : lo[0]);
            corners.push_back((k & 2) ? hi[1] : lo[1]);
            corners.push_back((k & 4) ? hi[2] : lo[2]);
        }
    } else {
        return set_error(MPM_ERR_INVALID, "unknown shape");
    }
    // world AABB (model.cpp:211-222) and range checks (MPMSimulator.cpp:210-222)
    double lb[3] = {1e300, 1e300, 1e300}, ub[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < corners.size(); i += 3) {
        double w[3];
        apply(M, &corners[i], w);
        for (int a = 0; a < 3; ++a) { lb[a] = std::min(lb[a], w[a]); ub[a] = std::max(ub[a], w[a]); }
    }
    const double h = cfg.h;
    const int nn[3] = {cfg.nx, cfg.ny, cfg.nz};
    for (int a = 0; a < 3; ++a) {
        double l = lb[a] - cfg.origin[a], u = ub[a] - cfg.origin[a];
        if (l < 0.0) return set_error(MPM_ERR_RANGE, "ERROR: OBJ lower bound is smaller than grid's origin.");
        if (u > (nn[a] - 1) * h) return set_error(MPM_ERR_RANGE, "ERROR: OBJ upper bound is out of the grid.");
    }
    int lo_i[3], hi_i[3];
    for (int a = 0; a < 3; ++a) {
        lo_i[a] = (int)((lb[a] - cfg.origin[a]) / h);          // :225-227
        hi_i[a] = (int)((ub[a] - cfg.origin[a]) / h) + 1;      // :228-230
        hi_i[a] = std::min(hi_i[a], nn[a] - 1);
    }
    const int ppc = std::max(1, cfg.ppc);
    const double r2 = obj.shape == MPM_SHAPE_SPHERE ? obj.shape_param[3] * obj.shape_param[3] : 0.0;
    uint64_t sample = 0;
    for (int ix = lo_i[0]; ix < hi_i[0]; ++ix)
        for (int iy = lo_i[1]; iy < hi_i[1]; ++iy)
            for (int iz = lo_i[2]; iz < hi_i[2]; ++iz)
                for (int ip = 0; ip < ppc; ++ip, ++sample) {
                    uint64_t id = (((uint64_t)ix * 65536u + (uint64_t)iy) * 65536u + (uint64_t)iz) * 1024u + (uint64_t)ip;
                    float p[3];
                    const int ii[3] = {ix, iy, iz};
                    for (int a = 0; a < 3; ++a) {
                        float r = unit_float(cfg.seed, id, (uint32_t)a);
                        p[a] = (float)(cfg.origin[a] + ii[a] * h + r * h);   // :246-248
                    }
                    double pw[3] = {p[0], p[1], p[2]}, pl[3];
                    apply(Minv, pw, pl);
                    bool in = false;
                    if (obj.shape == MPM_SHAPE_MESH) in = mesh.inside(pl);
                    else if (obj.shape == MPM_SHAPE_BOX)
                        in = pl[0] >= obj.shape_param[0] && pl[0] <= obj.shape_param[3] && pl[1] >= obj.shape_param[1] &&
                             pl[1] <= obj.shape_param[4] && pl[2] >= obj.shape_param[2] && pl[2] <= obj.shape_param[5];
                    else {
                        double dx = pl[0] - obj.shape_param[0], dy = pl[1] - obj.shape_param[1],
                               dz = pl[2] - obj.shape_param[2];
                        in = dx * dx + dy * dy + dz * dz <= r2;
                    }
                    if (in) { xyz.push_back(p[0]); xyz.push_back(p[1]); xyz.push_back(p[2]); }
                }
    int64_t n = (int64_t)xyz.size() / 3;
    if (obj.max_particles > 0 && n > obj.max_particles) {
        // seeded subset of exactly max_particles, original order kept
        std::vector<std::pair<uint64_t, int64_t>> key((size_t)n);
        for (int64_t i = 0; i < n; ++i) key[i] = {mix64(cfg.seed * 31u + 7u + mix64((uint64_t)i)), i};
        std::nth_element(key.begin(), key.begin() + obj.max_particles, key.end());
        std::vector<int64_t> keep((size_t)obj.max_particles);
        for (int64_t i = 0; i < obj.max_particles; ++i) keep[i] = key[i].second;
        std::sort(keep.begin(), keep.end());
        std::vector<float> out((size_t)obj.max_particles * 3);
        for (int64_t i = 0; i < obj.max_particles; ++i)
            for (int a = 0; a < 3; ++a) out[3 * i + a] = xyz[3 * keep[i] + a];
        xyz.swap(out);
    }
    return MPM_OK;
}

// ---------------------------------------------------------------------------
// Houdini classic BGEO, version 5: big-endian header (magic "Bgeo", 'V', 5,
// point / primitive / group / attribute counts), then per point x y z w, then
// the empty "extra" section markers 0x00 0xff.  One point attribute
// ("position", carried by the point coordinates themselves).
static void put_be32(std::vector<unsigned char> &b, uint32_t v)
{
    b.push_back((unsigned char)(v >> 24)); b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8)); b.push_back((unsigned char)v);
}

static void put_bef(std::vector<unsigned char> &b, float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    put_be32(b, u);
}

int write_bgeo(const char *path, const float *xyz, int64_t n)
{
    if (n > 0x7fffffff) return set_error(MPM_ERR_INVALID, "too many points for BGEO");
    std::vector<unsigned char> b;
    b.reserve(64 + (size_t)n * 16);
    put_be32(b, ((uint32_t)'B' << 24) | ((uint32_t)'g' << 16) | ((uint32_t)'e' << 8) | (uint32_t)'o');
    b.push_back((unsigned char)'V');
    put_be32(b, 5);                   // version
    put_be32(b, (uint32_t)n);         // points
    put_be32(b, 0);                   // primitives
    put_be32(b, 0);                   // point groups
    put_be32(b, 0);                   // primitive groups
    put_be32(b, 0);                   // point attributes besides position
    put_be32(b, 0);                   // vertex attributes
    put_be32(b, 0);                   // primitive attributes
    put_be32(b, 0);                   // detail attributes
    for (int64_t i = 0; i < n; ++i) {
        put_bef(b, xyz[3 * i]); put_bef(b, xyz[3 * i + 1]); put_bef(b, xyz[3 * i + 2]);
        put_bef(b, 1.0f);             // homogeneous w
    }
    b.push_back(0x00);                // begin extra
    b.push_back(0xff);                // end extra
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_error(MPM_ERR_IO, std::string("cannot open for writing: ") + path);
    size_t w = fwrite(b.data(), 1, b.size(), fp);
    int cl = fclose(fp);
    if (w != b.size() || cl != 0) return set_error(MPM_ERR_IO, std::string("short write: ") + path);
    return MPM_OK;
}

}  // namespace mpm
