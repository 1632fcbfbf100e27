// Synthetic pressure / momentum systems on hexahedral box meshes.
//
// The mesh follows the reference's structured box mesher (src/mesh/src/box_mesher.cpp:48-145):
// points i-fastest, cells k-outermost and i-innermost, internal faces ordered by
// (owner, neighbour) with owner < neighbour (src/mesh/src/mesh_builder.cpp:155-171). Face
// geometry restates Mesh::compute_geometry (src/mesh/src/mesh.cpp:100-186): face centre and
// area vector from a triangle fan about the vertex average, cell centre and volume from a
// pyramid decomposition about the average of the face centres. Interior vertices can be
// jittered (non-orthogonal, non-planar faces). Per internal face the generator returns the
// coefficient |S_f| / |C_N - C_P| (tests/test_amg.cpp:52-61) and the linear interpolation
// weight w = S.(C_f - C_P) / (S.(C_f - C_P) + S.(C_N - C_f)) (src/fv/src/fv_geometry.cpp:52-58).
#include "common.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <random>

namespace cfdg {
namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

constexpr double k_tiny = 1e-300; // src/mesh/src/mesh.cpp:11

double unit_draw(std::mt19937_64& g) {
    return static_cast<double>(g() >> 11) * 0x1.0p-53;
}

struct Box {
    int nx, ny, nz;
    std::vector<V3> pts; // (nx+1)(ny+1)(nz+1), i fastest

    [[nodiscard]] long long pid(int i, int j, int k) const {
        return i + static_cast<long long>(nx + 1) * (j + static_cast<long long>(ny + 1) * k);
    }
    [[nodiscard]] long long cid(int i, int j, int k) const {
        return i + static_cast<long long>(nx) * (j + static_cast<long long>(ny) * k);
    }
};

void check_spec(const cfdg_box_spec* s) {
    require(s != nullptr, "box spec is null");
    require(s->nx >= 1 && s->ny >= 1 && s->nz >= 1, "make_box_mesh: cell counts must be >= 1");
    for (int d = 0; d < 3; ++d) {
        require(s->hi[d] > s->lo[d], "make_box_mesh: max must be greater than min");
    }
    require(s->jitter >= 0.0 && s->jitter < 0.5, "jitter must be in [0, 0.5)");
    const long long nc = static_cast<long long>(s->nx) * s->ny * s->nz;
    const long long nif = static_cast<long long>(s->nx - 1) * s->ny * s->nz +
                          static_cast<long long>(s->nx) * (s->ny - 1) * s->nz +
                          static_cast<long long>(s->nx) * s->ny * (s->nz - 1);
    require(nc < (1LL << 31) && nc + 2 * nif < (1LL << 31), "box too large for int32 indices");
}

Box make_points(const cfdg_box_spec& s) {
    Box b{s.nx, s.ny, s.nz, {}};
    b.pts.resize(static_cast<std::size_t>(s.nx + 1) * (s.ny + 1) * (s.nz + 1));
    const double hx = (s.hi[0] - s.lo[0]) / s.nx;
    const double hy = (s.hi[1] - s.lo[1]) / s.ny;
    const double hz = (s.hi[2] - s.lo[2]) / s.nz;
    const double ex = s.hi[0] - s.lo[0];
    const double ey = s.hi[1] - s.lo[1];
    const double ez = s.hi[2] - s.lo[2];
    std::mt19937_64 gen(s.seed);
    for (int k = 0; k <= s.nz; ++k) {
        for (int j = 0; j <= s.ny; ++j) {
            for (int i = 0; i <= s.nx; ++i) {
                // Uniform node fractions i/n as graded_positions with grading 1
                // (box_mesher.cpp:36-46, 62-64).
                V3 p{s.lo[0] + ex * (static_cast<double>(i) / s.nx),
                     s.lo[1] + ey * (static_cast<double>(j) / s.ny),
                     s.lo[2] + ez * (static_cast<double>(k) / s.nz)};
                const bool interior =
                    i > 0 && i < s.nx && j > 0 && j < s.ny && k > 0 && k < s.nz;
                if (interior && s.jitter > 0.0) {
                    const double ux = unit_draw(gen);
                    const double uy = unit_draw(gen);
                    const double uz = unit_draw(gen);
                    p.x += (2.0 * ux - 1.0) * s.jitter * hx;
                    p.y += (2.0 * uy - 1.0) * s.jitter * hy;
                    p.z += (2.0 * uz - 1.0) * s.jitter * hz;
                }
                b.pts[static_cast<std::size_t>(b.pid(i, j, k))] = p;
            }
        }
    }
    return b;
}

struct FaceGeo {
    V3 centre;
    V3 area; // oriented towards +axis
};

// Quad face geometry by a triangle fan about the vertex average (mesh.cpp:118-137).
FaceGeo quad_geometry(const std::array<V3, 4>& p) {
    V3 pc{0.0, 0.0, 0.0};
    for (const V3& q : p) {
        pc = pc + q;
    }
    pc = pc / 4.0;
    V3 sum_n{0.0, 0.0, 0.0};
    V3 sum_ac{0.0, 0.0, 0.0};
    double sum_a = 0.0;
    for (int i = 0; i < 4; ++i) {
        const V3& p0 = p[static_cast<std::size_t>(i)];
        const V3& p1 = p[static_cast<std::size_t>((i + 1) % 4)];
        const V3 n = cross(p1 - p0, pc - p0);
        const double a = norm(n);
        sum_n = sum_n + n;
        sum_a += a;
        sum_ac = sum_ac + a * ((p0 + p1) + pc);
    }
    FaceGeo g;
    g.centre = sum_a > k_tiny ? sum_ac / (3.0 * sum_a) : pc;
    g.area = 0.5 * sum_n;
    return g;
}

// Face normal to axis `ax` at lattice position (i, j, k) of its lowest corner, with the node
// order chosen so that the area vector points towards +axis.
FaceGeo face_geo(const Box& b, int ax, int i, int j, int k) {
    std::array<long long, 4> id{};
    if (ax == 0) { // nodes in the y-z plane: (j,k) (j+1,k) (j+1,k+1) (j,k+1)
        id = {b.pid(i, j, k), b.pid(i, j + 1, k), b.pid(i, j + 1, k + 1), b.pid(i, j, k + 1)};
    } else if (ax == 1) { // z then x: (k,i) (k+1,i) (k+1,i+1) (k,i+1)
        id = {b.pid(i, j, k), b.pid(i, j, k + 1), b.pid(i + 1, j, k + 1), b.pid(i + 1, j, k)};
    } else { // x then y
        id = {b.pid(i, j, k), b.pid(i + 1, j, k), b.pid(i + 1, j + 1, k), b.pid(i, j + 1, k)};
    }
    std::array<V3, 4> p{};
    for (int q = 0; q < 4; ++q) {
        p[static_cast<std::size_t>(q)] = b.pts[static_cast<std::size_t>(id[static_cast<std::size_t>(q)])];
    }
    return quad_geometry(p);
}

struct CellGeo {
    V3 centre;
    double volume;
};

// Pyramid decomposition of one hex cell (mesh.cpp:140-185). Faces are visited in the
// reference's global face order: internal faces by (owner, neighbour) – the three faces the
// cell receives from lower-index owners (z, y, x neighbour below), then the faces it owns
// (x, y, z neighbour above) – followed by boundary faces in patch order xmin, xmax, ymin,
// ymax, zmin, zmax.
CellGeo cell_geometry(const Box& b, int i, int j, int k) {
    struct F {
        int key0;     // 0 internal, 1 boundary
        long long k1; // owner or patch
        long long k2; // neighbour
        FaceGeo g;
        double sign; // +1: outward along +axis
    };
    std::array<F, 6> f{};
    const long long c = b.cid(i, j, k);
    int m = 0;
    auto add = [&](int ax, int fi, int fj, int fk, bool max_side, bool boundary, long long nb,
                   int patch) {
        F e;
        e.g = face_geo(b, ax, fi, fj, fk);
        e.sign = max_side ? 1.0 : -1.0;
        if (boundary) {
            e.key0 = 1;
            e.k1 = patch;
            e.k2 = 0;
        } else {
            e.key0 = 0;
            e.k1 = std::min(c, nb);
            e.k2 = std::max(c, nb);
        }
        f[static_cast<std::size_t>(m++)] = e;
    };
    add(0, i, j, k, false, i == 0, i > 0 ? b.cid(i - 1, j, k) : -1, 0);
    add(0, i + 1, j, k, true, i == b.nx - 1, i < b.nx - 1 ? b.cid(i + 1, j, k) : -1, 1);
    add(1, i, j, k, false, j == 0, j > 0 ? b.cid(i, j - 1, k) : -1, 2);
    add(1, i, j + 1, k, true, j == b.ny - 1, j < b.ny - 1 ? b.cid(i, j + 1, k) : -1, 3);
    add(2, i, j, k, false, k == 0, k > 0 ? b.cid(i, j, k - 1) : -1, 4);
    add(2, i, j, k + 1, true, k == b.nz - 1, k < b.nz - 1 ? b.cid(i, j, k + 1) : -1, 5);
    std::sort(f.begin(), f.end(), [](const F& a, const F& b2) {
        if (a.key0 != b2.key0) {
            return a.key0 < b2.key0;
        }
        if (a.k1 != b2.k1) {
            return a.k1 < b2.k1;
        }
        return a.k2 < b2.k2;
    });
    V3 ce{0.0, 0.0, 0.0};
    for (const F& e : f) {
        ce = ce + e.g.centre;
    }
    ce = ce / 6.0;
    double pyr3 = 0.0;
    V3 vc{0.0, 0.0, 0.0};
    for (const F& e : f) {
        // Outward area vector dotted with (C_f - C_e): equals the reference's owner term
        // S.(C_f - C_e) and neighbour term S.(C_e - C_f) with S pointing into the neighbour.
        const V3 s = e.sign * e.g.area;
        const double v = dot(s, e.g.centre - ce);
        pyr3 += v;
        vc = vc + v * (0.75 * e.g.centre + 0.25 * ce);
    }
    CellGeo g;
    g.volume = pyr3 / 3.0;
    g.centre = std::abs(pyr3) > k_tiny ? vc / pyr3 : ce;
    return g;
}

} // namespace

} // namespace cfdg

using namespace cfdg;

extern "C" cfdg_status cfdg_box_sizes(const cfdg_box_spec* spec, int* n_cells,
                                      int* n_internal_faces) {
    return guarded([&] {
        check_spec(spec);
        require(n_cells != nullptr && n_internal_faces != nullptr, "null output");
        *n_cells = spec->nx * spec->ny * spec->nz;
        *n_internal_faces = (spec->nx - 1) * spec->ny * spec->nz +
                            spec->nx * (spec->ny - 1) * spec->nz +
                            spec->nx * spec->ny * (spec->nz - 1);
    });
}

extern "C" cfdg_status cfdg_box_build(const cfdg_box_spec* spec, int* owner, int* neighbour,
                                      double* area_over_distance, double* weight,
                                      double* cell_volume, double* cell_centre) {
    return guarded([&] {
        check_spec(spec);
        require(owner != nullptr && neighbour != nullptr, "owner / neighbour must not be null");
        const Box b = make_points(*spec);
        const int nx = b.nx;
        const int ny = b.ny;
        const int nz = b.nz;
        const long long nc = static_cast<long long>(nx) * ny * nz;
        std::vector<V3> cc(static_cast<std::size_t>(nc));
#pragma omp parallel for schedule(static)
        for (long long c = 0; c < nc; ++c) {
            const int i = static_cast<int>(c % nx);
            const int j = static_cast<int>((c / nx) % ny);
            const int k = static_cast<int>(c / (static_cast<long long>(nx) * ny));
            const CellGeo g = cell_geometry(b, i, j, k);
            cc[static_cast<std::size_t>(c)] = g.centre;
            if (cell_volume != nullptr) {
                cell_volume[c] = g.volume;
            }
            if (cell_centre != nullptr) {
                cell_centre[3 * c + 0] = g.centre.x;
                cell_centre[3 * c + 1] = g.centre.y;
                cell_centre[3 * c + 2] = g.centre.z;
            }
        }
        // Internal faces of cell c in (owner, neighbour) order: c+1, c+nx, c+nx*ny. The face
        // index of the first face owned by c is the number of internal faces owned by cells
        // below c, which follows from counting per row of the lattice.
        std::vector<long long> first(static_cast<std::size_t>(nc) + 1, 0);
        for (long long c = 0; c < nc; ++c) {
            const int i = static_cast<int>(c % nx);
            const int j = static_cast<int>((c / nx) % ny);
            const int k = static_cast<int>(c / (static_cast<long long>(nx) * ny));
            first[static_cast<std::size_t>(c) + 1] = first[static_cast<std::size_t>(c)] +
                                                      (i < nx - 1) + (j < ny - 1) + (k < nz - 1);
        }
#pragma omp parallel for schedule(static)
        for (long long c = 0; c < nc; ++c) {
            const int i = static_cast<int>(c % nx);
            const int j = static_cast<int>((c / nx) % ny);
            const int k = static_cast<int>(c / (static_cast<long long>(nx) * ny));
            long long f = first[static_cast<std::size_t>(c)];
            const V3 cp = cc[static_cast<std::size_t>(c)];
            auto emit = [&](int ax, int fi, int fj, int fk, long long nb) {
                const FaceGeo g = face_geo(b, ax, fi, fj, fk);
                const V3 cn = cc[static_cast<std::size_t>(nb)];
                owner[f] = static_cast<int>(c);
                neighbour[f] = static_cast<int>(nb);
                if (area_over_distance != nullptr) {
                    area_over_distance[f] = norm(g.area) / norm(cn - cp);
                }
                if (weight != nullptr) {
                    const double dp = dot(g.area, g.centre - cp);
                    const double dn = dot(g.area, cn - g.centre);
                    const double sum = dp + dn;
                    weight[f] = sum > 0.0 ? std::clamp(dn / sum, 0.0, 1.0) : 0.5;
                }
                ++f;
            };
            if (i < nx - 1) {
                emit(0, i + 1, j, k, c + 1);
            }
            if (j < ny - 1) {
                emit(1, i, j + 1, k, c + nx);
            }
            if (k < nz - 1) {
                emit(2, i, j, k + 1, c + static_cast<long long>(nx) * ny);
            }
        }
    });
}

extern "C" cfdg_status cfdg_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    return guarded([&] {
        require(n >= 0 && (n == 0 || out != nullptr), "invalid uniform request");
        std::mt19937_64 gen(seed);
        const double span = hi - lo;
        for (int64_t i = 0; i < n; ++i) {
            out[i] = lo + span * unit_draw(gen);
        }
    });
}

extern "C" cfdg_status cfdg_csr_pattern(int n, int n_faces, const int* owner,
                                        const int* neighbour, int* row_ptr, int* col,
                                        int* diag_pos, int* upper_pos, int* lower_pos) {
    return guarded([&] {
        require(n >= 0 && n_faces >= 0, "negative size");
        require(row_ptr != nullptr && col != nullptr, "null output");
        require(n_faces == 0 || (owner != nullptr && neighbour != nullptr), "null face arrays");
        for (int f = 0; f < n_faces; ++f) {
            require(owner[f] >= 0 && owner[f] < n && neighbour[f] >= 0 && neighbour[f] < n &&
                        owner[f] != neighbour[f],
                    "face cell index out of range");
        }
        // Row lengths: the diagonal plus one entry per adjacent face; the mesh invariants
        // (one face per cell pair) make setFromTriplets' duplicate summation a no-op.
        std::vector<int> len(static_cast<std::size_t>(n), 1);
        for (int f = 0; f < n_faces; ++f) {
            ++len[static_cast<std::size_t>(owner[f])];
            ++len[static_cast<std::size_t>(neighbour[f])];
        }
        row_ptr[0] = 0;
        for (int i = 0; i < n; ++i) {
            row_ptr[i + 1] = row_ptr[i] + len[static_cast<std::size_t>(i)];
        }
        std::vector<int> fill(row_ptr, row_ptr + n);
        for (int i = 0; i < n; ++i) {
            col[fill[static_cast<std::size_t>(i)]++] = i;
        }
        for (int f = 0; f < n_faces; ++f) {
            col[fill[static_cast<std::size_t>(owner[f])]++] = neighbour[f];
            col[fill[static_cast<std::size_t>(neighbour[f])]++] = owner[f];
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            std::sort(col + row_ptr[i], col + row_ptr[i + 1]);
        }
        for (int i = 0; i < n; ++i) {
            for (int k = row_ptr[i] + 1; k < row_ptr[i + 1]; ++k) {
                require(col[k] != col[k - 1], "duplicate face between the same two cells");
            }
        }
        auto find = [&](int r, int c) {
            const int* b = col + row_ptr[r];
            const int* e = col + row_ptr[r + 1];
            const int* it = std::lower_bound(b, e, c);
            return static_cast<int>(it - col);
        };
        if (diag_pos != nullptr) {
            for (int i = 0; i < n; ++i) {
                diag_pos[i] = find(i, i);
            }
        }
        for (int f = 0; f < n_faces; ++f) {
            if (upper_pos != nullptr) {
                upper_pos[f] = find(owner[f], neighbour[f]);
            }
            if (lower_pos != nullptr) {
                lower_pos[f] = find(neighbour[f], owner[f]);
            }
        }
    });
}
