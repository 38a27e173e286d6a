// Host-side sanitizer check (AddressSanitizer + UndefinedBehaviorSanitizer
// on the host code of libflgpu and on the oracle; device code is built as
// usual).  Built by tests/sanitize/Makefile, run by
// tests/test_host_sanitizers.py on the CPU:
//   - make_sparsity and colour_mesh on a jittered-free Kuhn cube mesh;
//   - the sliced ELL plan of the solver's SpMV against its CSR source;
//   - the matrix-dump writer and reader, then the reader on every
//     truncation of a dump and on dumps with corrupted headers, counts,
//     row lengths and column indices (each must fail cleanly or succeed);
//   - the oracle's assembly, Dirichlet, solves (CG / GMRES, Jacobi / SSOR),
//     colourings and pressure path on the same mesh.
// With --gpu it also binds the mesh to a device context and runs the
// blocked assembly (the host-side slab plan builder), the two-pass P2
// plan, the solve and the pressure path.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../fluidity_amd/csrc/krylov_common.h"
#include "../../include/flgpu.h"
#include "../../oracle/fl_oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::fprintf(stderr, "CHECK failed: " __VA_ARGS__); \
            std::fprintf(stderr, " (%s:%d)\n", __FILE__, __LINE__); \
            failures++;                                   \
        }                                                 \
    } while (0)

// n^3 cubes, each split into 6 tets around the main diagonal (1-based).
struct Mesh {
    int nx = 0, nele = 0;
    std::vector<double> X;       // 3 x nx
    std::vector<int32_t> ndglno; // 4 x nele
};

static Mesh kuhn_cube(int n) {
    Mesh m;
    const int p = n + 1;
    m.nx = p * p * p;
    m.X.resize(3 * (size_t)m.nx);
    for (int k = 0; k < p; k++)
        for (int j = 0; j < p; j++)
            for (int i = 0; i < p; i++) {
                const int v = i + p * (j + p * k);
                m.X[3 * v] = (double)i / n;
                m.X[3 * v + 1] = (double)j / n;
                m.X[3 * v + 2] = (double)k / n;
            }
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                for (int t = 0; t < 6; t++) {
                    int c[3] = {i, j, k};
                    int tet[4];
                    tet[0] = c[0] + p * (c[1] + p * c[2]);
                    for (int s = 0; s < 3; s++) {
                        c[perm[t][s]]++;
                        tet[s + 1] = c[0] + p * (c[1] + p * c[2]);
                    }
                    for (int s = 0; s < 4; s++) m.ndglno.push_back(tet[s] + 1);
                }
    m.nele = (int)m.ndglno.size() / 4;
    return m;
}

static std::string tmp_path(const char* name) {
    const char* d = std::getenv("TMPDIR");
    return std::string(d ? d : "/tmp") + "/" + name;
}

static void write_bytes(const std::string& path, const std::vector<unsigned char>& b) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

static std::vector<unsigned char> read_bytes(const std::string& path) {
    std::vector<unsigned char> b;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return b;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) b.push_back((unsigned char)ch);
    std::fclose(f);
    return b;
}

// Read a dump the way a caller does: size query, then arrays of that size.
static int read_dump(const std::string& path, int32_t* n_out) {
    int32_t n = 0, has_x0 = 0;
    int64_t nnz = 0;
    int rc = flgpu_matrixdump_read(path.c_str(), &n, &nnz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc != FLGPU_OK) return rc;
    if (n < 0 || nnz < 0 || n > (1 << 24) || nnz > (1 << 26)) return -100;  // the query bounds what is allocated
    std::vector<int32_t> f((size_t)n + 1), c((size_t)nnz);
    std::vector<double> v((size_t)nnz), b((size_t)n), x0((size_t)n);
    rc = flgpu_matrixdump_read(path.c_str(), &n, &nnz, f.data(), c.data(), v.data(), b.data(), x0.data(), &has_x0);
    if (n_out) *n_out = n;
    return rc;
}

static void check_matrixdump(const std::vector<int32_t>& findrm, const std::vector<int32_t>& colm) {
    const int n = (int)findrm.size() - 1;
    std::vector<double> val(colm.size()), b(n), x0(n);
    for (size_t k = 0; k < val.size(); k++) val[k] = std::sin(0.1 * (double)k) + 2.0;
    for (int i = 0; i < n; i++) b[i] = i, x0[i] = -i;
    const std::string path = tmp_path("flgpu_host_check.dump");
    CHECK(flgpu_matrixdump_write(path.c_str(), n, findrm.data(), colm.data(), val.data(), b.data(), x0.data()) ==
              FLGPU_OK, "matrixdump write");
    int32_t n2 = 0;
    CHECK(read_dump(path, &n2) == FLGPU_OK && n2 == n, "matrixdump read back");
    const std::vector<unsigned char> good = read_bytes(path);
    const std::string bad = tmp_path("flgpu_host_check_bad.dump");
    int ok = 0, refused = 0;
    // every truncation of the first 4 KiB and a spread of longer ones
    for (size_t len = 0; len < good.size(); len += (len < 4096 ? 1 : 997)) {
        write_bytes(bad, std::vector<unsigned char>(good.begin(), good.begin() + len));
        (read_dump(bad, nullptr) == FLGPU_OK ? ok : refused)++;
    }
    // corrupted words: headers, counts, row lengths, column indices, values
    std::mt19937_64 rng(12345);
    const uint32_t specials[] = {0u, 1u, 0x7fffffffu, 0x80000000u, 0xffffffffu, 1211216u, 1211214u, 3u};
    for (int trial = 0; trial < 3000; trial++) {
        std::vector<unsigned char> b2 = good;
        const int nflip = 1 + (int)(rng() % 3);
        for (int q = 0; q < nflip; q++) {
            size_t at;
            if (trial % 3 == 0) at = (rng() % 8) * 4;                      // a header word
            else at = (rng() % (b2.size() / 4)) * 4;                        // any word
            const uint32_t w = (trial % 2) ? specials[rng() % 8] : (uint32_t)rng();
            for (int s = 0; s < 4 && at + s < b2.size(); s++) b2[at + s] = (unsigned char)(w >> (24 - 8 * s));
        }
        write_bytes(bad, b2);
        (read_dump(bad, nullptr) == FLGPU_OK ? ok : refused)++;
    }
    std::printf("matrixdump fuzz: %d read, %d refused\n", ok, refused);
    CHECK(flgpu_matrixdump_read(tmp_path("flgpu_no_such_file").c_str(), &n2, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr) != FLGPU_OK, "missing file");
    std::remove(path.c_str());
    std::remove(bad.c_str());
}

static void check_oracle(const Mesh& m, const std::vector<int32_t>& findrm, const std::vector<int32_t>& colm,
                         const std::vector<int32_t>& centrm) {
    double l[64 * 4], w[64], n[64 * 10], dn[64 * 10 * 3];
    const int ngi = orc_quadrature(3, 2, l, w);
    CHECK(ngi == 4, "quadrature");
    CHECK(orc_shape_tables(3, 1, ngi, l, n, dn) == 4, "shape tables");
    std::vector<int32_t> colour(m.nele);
    const int ncol = orc_colour_elements(m.nx, m.nele, 4, m.ndglno.data(), colour.data());
    std::vector<int32_t> col_ptr(ncol + 1, 0), elems(m.nele);
    for (int e = 0; e < m.nele; e++) col_ptr[colour[e]]++;
    for (int c = 0; c < ncol; c++) col_ptr[c + 1] += col_ptr[c];
    std::vector<int32_t> fill(col_ptr.begin(), col_ptr.end() - 1);
    for (int e = 0; e < m.nele; e++) elems[fill[colour[e] - 1]++] = e + 1;
    std::vector<double> T(m.nx), kappa(9, 0.0), src(m.nx);
    kappa[0] = kappa[4] = kappa[8] = 1.0;
    for (int i = 0; i < m.nx; i++) T[i] = std::sin(3.0 * m.X[3 * i]), src[i] = 1.0;
    orc_mesh om{};
    om.dim = 3;
    om.nele = m.nele;
    om.loc_x = 4;
    om.nx = m.nx;
    om.ndglno_x = m.ndglno.data();
    om.X = m.X.data();
    om.x_dn = dn;
    om.loc = 4;
    om.ngi = ngi;
    om.nnodes = m.nx;
    om.ndglno = m.ndglno.data();
    om.n = n;
    om.dn = dn;
    om.w = w;
    om.T = T.data();
    om.kappa = kappa.data();
    om.kappa_const = 1;
    om.src = src.data();
    om.ncolours = ncol;
    om.col_ptr = col_ptr.data();
    om.col_elems = elems.data();
    om.nrows = m.nx;
    om.findrm = findrm.data();
    om.colm = colm.data();
    orc_params p{};
    p.dt = 0.1;
    p.theta = 0.5;
    p.have_mass = 1;
    p.have_diffusivity = 1;
    p.isotropic_diffusivity = 1;
    p.have_source = 1;
    std::vector<double> A(colm.size()), rhs(m.nx);
    CHECK(orc_assemble(&p, &om, A.data(), rhs.data(), 1, nullptr) == 0, "oracle assemble");
    std::vector<int32_t> inactive(m.nx, 0), bnodes;
    for (int i = 0; i < m.nx; i++)
        for (int a = 0; a < 3; a++)
            if (m.X[3 * i + a] == 0.0 || m.X[3 * i + a] == 1.0) {
                bnodes.push_back(i + 1);
                break;
            }
    std::vector<double> zero(bnodes.size(), 0.0);
    orc_dirichlet((int)bnodes.size(), bnodes.data(), zero.data(), T.data(), p.dt, inactive.data(), rhs.data());
    for (int ksp = 0; ksp < 2; ksp++)
        for (int pc : {1, 3}) {
            orc_solver s{ksp, pc, 1e-10, 0.0, 2000, 30, 1};
            std::vector<double> x(m.nx, 0.0);
            int32_t its = 0, reason = 0;
            double relres = 0.0;
            orc_petsc_solve(&s, m.nx, findrm.data(), colm.data(), centrm.data(), A.data(), inactive.data(),
                            rhs.data(), x.data(), &its, &reason, &relres, nullptr);
            CHECK(reason > 0, "oracle solve ksp %d pc %d", ksp, pc);
        }
    // the pressure path on the same mesh (P1-P1)
    double one = 1.0;
    orc_pressure pp{};
    pp.p_loc = 4;
    pp.p_nodes = m.nx;
    pp.p_ndglno = m.ndglno.data();
    pp.p_n = n;
    pp.p_dn = dn;
    pp.ct_findrm = findrm.data();
    pp.ct_colm = colm.data();
    pp.density = &one;
    pp.density_const = 1;
    std::vector<double> ct(3 * colm.size()), ml(m.nx);
    CHECK(orc_ct_assemble(&om, &pp, ct.data(), ml.data()) == 0, "oracle ct");
    std::vector<double> inv(3 * (size_t)m.nx);
    for (int i = 0; i < m.nx; i++)
        for (int a = 0; a < 3; a++) inv[a + 3 * (size_t)i] = 1.0 / ml[i];
    std::vector<double> cmc(colm.size());  // the first-order pattern, a subset of the product's
    orc_cmc(m.nx, findrm.data(), colm.data(), ct.data(), (int64_t)colm.size(), findrm.data(), colm.data(), ct.data(),
            (int64_t)colm.size(), 3, inv.data(), findrm.data(), colm.data(), cmc.data());
    std::printf("oracle: %d colours, assembly, 4 solves, pressure path\n", ncol);
}

static void check_gpu(const Mesh& m, const std::vector<int32_t>& findrm, const std::vector<int32_t>& colm,
                      const std::vector<int32_t>& centrm) {
    flgpu_ctx* ctx = nullptr;
    if (flgpu_init(&ctx, 0) != FLGPU_OK) {
        std::printf("gpu: %s\n", flgpu_last_error());
        failures++;
        return;
    }
    double l[64 * 4], w[64], n[64 * 10], dn[64 * 10 * 3];
    const int ngi = orc_quadrature(3, 2, l, w);
    orc_shape_tables(3, 1, ngi, l, n, dn);
    CHECK(flgpu_mesh_bind(ctx, 1, 0, 3, 4, m.nele, m.nx, m.ndglno.data(), m.X.data()) == FLGPU_OK, "%s",
          flgpu_last_error());
    CHECK(flgpu_field_mesh_bind(ctx, 4, ngi, m.nx, m.ndglno.data(), n, dn, w, dn) == FLGPU_OK, "%s",
          flgpu_last_error());
    CHECK(flgpu_sparsity_bind(ctx, m.nx, m.nx, findrm.data(), colm.data(), centrm.data()) == FLGPU_OK, "%s",
          flgpu_last_error());
    int32_t ncol = 0;
    CHECK(flgpu_colouring_build(ctx, &ncol) == FLGPU_OK, "%s", flgpu_last_error());
    std::vector<double> T(m.nx, 0.0);
    double kap[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, src = 1.0;
    CHECK(flgpu_field_set(ctx, FLGPU_FIELD_T, 0, T.data()) == FLGPU_OK, "%s", flgpu_last_error());
    CHECK(flgpu_field_set(ctx, FLGPU_FIELD_DIFFUSIVITY, 1, kap) == FLGPU_OK, "%s", flgpu_last_error());
    CHECK(flgpu_field_set(ctx, FLGPU_FIELD_SOURCE, 1, &src) == FLGPU_OK, "%s", flgpu_last_error());
    flgpu_advdif_params p{};
    p.dt = 1.0;
    p.theta = 1.0;
    p.have_diffusivity = 1;
    p.isotropic_diffusivity = 1;
    p.have_source = 1;
    p.scheme = FLGPU_ASM_BLOCKED;
    CHECK(flgpu_advdif_assemble_resident(ctx, &p) == FLGPU_OK, "blocked: %s", flgpu_last_error());
    p.scheme = FLGPU_ASM_COLOUR;
    CHECK(flgpu_advdif_assemble_resident(ctx, &p) == FLGPU_OK, "colour: %s", flgpu_last_error());
    std::vector<int32_t> bnodes;
    for (int i = 0; i < m.nx; i++)
        for (int a = 0; a < 3; a++)
            if (m.X[3 * i + a] == 0.0 || m.X[3 * i + a] == 1.0) {
                bnodes.push_back(i + 1);
                break;
            }
    std::vector<double> zero(bnodes.size(), 0.0);
    CHECK(flgpu_dirichlet_apply(ctx, (int)bnodes.size(), bnodes.data(), zero.data(), 0.0) == FLGPU_OK, "%s",
          flgpu_last_error());
    flgpu_solver_params sp{};
    sp.ksp = FLGPU_KSP_CG;
    sp.pc = FLGPU_PC_JACOBI;
    sp.rtol = 1e-10;
    sp.max_its = 2000;
    sp.start_from_zero = 1;
    sp.norm_type = FLGPU_NORM_UNPRECONDITIONED;
    std::vector<double> x(m.nx, 0.0);
    flgpu_solve_info info{};
    CHECK(flgpu_solve(ctx, &sp, x.data(), &info) == FLGPU_OK && info.reason > 0, "solve: %s", flgpu_last_error());
    CHECK(flgpu_solve_noalloc(ctx, &sp, x.data(), &info) == FLGPU_OK && info.reason > 0, "noalloc: %s",
          flgpu_last_error());
    CHECK(flgpu_pressure_bind(ctx, 4, m.nx, m.ndglno.data(), n, dn, m.nx, findrm.data(), colm.data()) == FLGPU_OK,
          "%s", flgpu_last_error());
    double one = 1.0;
    CHECK(flgpu_ct_assemble(ctx, 0, 1, &one, nullptr, nullptr) == FLGPU_OK, "%s", flgpu_last_error());
    std::vector<double> cmc(colm.size());
    CHECK(flgpu_cmc_assemble(ctx, nullptr, m.nx, findrm.data(), colm.data(), cmc.data()) == FLGPU_OK, "%s",
          flgpu_last_error());
    CHECK(flgpu_finalize(ctx) == FLGPU_OK, "finalize");
    std::printf("gpu: blocked and colour assembly, solve, no-alloc solve, pressure path (%d its)\n", info.its);
}

// The sliced ELL plan (flgpu::sell_plan) against its CSR source, walked
// with the GPU kernel's index arithmetic: every entry in exactly one slot,
// at slot sptr[s] + 64 j + lane of its row, its column recovered from the
// stored index, padding marked -1 with an in-range column; the product in
// slot order equals the CSR product bit for bit.
static void check_sell(const std::vector<int32_t>& findrm, const std::vector<int32_t>& colm) {
    const int n = (int)findrm.size() - 1;
    std::vector<int32_t> rp(n + 1), cl(colm.size()), vmap(colm.size());
    for (int i = 0; i <= n; i++) rp[i] = findrm[i] - 1;
    for (size_t k = 0; k < colm.size(); k++) {
        cl[k] = colm[k] - 1;
        vmap[k] = 3 * (int32_t)k + 1;
    }
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> val(cl.size()), x(n), ycsr(n), ysell(n);
    for (double& v : val) v = U(g);
    for (double& v : x) v = U(g);
    for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) acc += val[k] * x[cl[k]];
        ycsr[i] = acc;
    }
    for (int use_map = 0; use_map < 2; use_map++) {
        flgpu::SellHost h;
        flgpu::sell_plan(rp.data(), cl.data(), n, use_map ? vmap.data() : nullptr, h);
        CHECK(h.nsl == (n + flgpu::kSlice - 1) / flgpu::kSlice, "sell slices");
        CHECK(h.nslot == h.sptr[h.nsl] && (int64_t)h.src.size() == h.nslot, "sell slots");
        std::vector<int> seen(cl.size(), 0);
        for (int sl = 0; sl < h.nsl; sl++) {
            const int r0 = sl * flgpu::kSlice, base = h.sptr[sl];
            CHECK((h.sptr[sl + 1] - base) % flgpu::kSlice == 0, "sell slice width");
            const int w = (h.sptr[sl + 1] - base) / flgpu::kSlice;
            for (int lane = 0; lane < flgpu::kSlice; lane++) {
                const int row = r0 + lane;
                const int len = row < n ? rp[row + 1] - rp[row] : 0;
                CHECK(len <= w, "sell row longer than its slice");
                double acc = 0.0;
                for (int j = 0; j < w; j++) {
                    const int64_t slot = base + (int64_t)j * flgpu::kSlice + lane;
                    const int cj = h.use16 ? r0 + h.c16[slot] : h.c32[slot];
                    CHECK(cj >= 0 && cj < n, "sell column out of range");
                    if (j < len) {
                        const int k = rp[row] + j;
                        CHECK(h.src[slot] == (use_map ? vmap[k] : k), "sell source entry");
                        CHECK(cj == cl[k], "sell column");
                        seen[k]++;
                        acc += val[k] * x[cj];
                    } else {
                        CHECK(h.src[slot] == -1, "sell padding");
                    }
                }
                if (row < n) ysell[row] = acc;
            }
        }
        for (size_t k = 0; k < seen.size(); k++) CHECK(seen[k] == 1, "sell entry %zu in %d slots\n", k, seen[k]);
        CHECK(std::memcmp(ycsr.data(), ysell.data(), sizeof(double) * n) == 0, "sell product differs from CSR");
        std::printf("sell: %d slices, %lld slots for %zu entries (fill %.3f), int16 columns %d\n", h.nsl,
                    (long long)h.nslot, cl.size(), h.fill((int64_t)cl.size()), h.use16 ? 1 : 0);
    }
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    const Mesh m = kuhn_cube(gpu ? 12 : 6);
    std::vector<int32_t> findrm(m.nx + 1), centrm(m.nx);
    int64_t nnz = 0;
    CHECK(flgpu_make_sparsity(m.nx, m.nele, 4, m.ndglno.data(), findrm.data(), nullptr, centrm.data(), &nnz) ==
              FLGPU_OK, "make_sparsity size");
    std::vector<int32_t> colm((size_t)nnz);
    CHECK(flgpu_make_sparsity(m.nx, m.nele, 4, m.ndglno.data(), findrm.data(), colm.data(), centrm.data(), &nnz) ==
              FLGPU_OK, "make_sparsity");
    std::vector<int32_t> colour(m.nele);
    int32_t ncol = 0;
    CHECK(flgpu_colour_mesh(m.nx, m.nele, 4, m.ndglno.data(), colour.data(), &ncol) == FLGPU_OK, "colour_mesh");
    // corrupt inputs to the host helpers fail cleanly
    std::vector<int32_t> badconn = m.ndglno;
    badconn[5] = m.nx + 7;
    CHECK(flgpu_make_sparsity(m.nx, m.nele, 4, badconn.data(), findrm.data(), nullptr, centrm.data(), &nnz) !=
              FLGPU_OK, "bad ndglno refused");
    CHECK(flgpu_make_sparsity(m.nx, m.nele, 4, m.ndglno.data(), findrm.data(), colm.data(), centrm.data(), &nnz) ==
              FLGPU_OK, "make_sparsity again");
    std::printf("host: %d nodes, %d elements, %lld entries, %d colours\n", m.nx, m.nele, (long long)nnz, ncol);
    check_matrixdump(findrm, colm);
    check_sell(findrm, colm);
    check_oracle(m, findrm, colm, centrm);
    if (gpu) check_gpu(m, findrm, colm, centrm);
    std::printf("%s\n", failures ? "FAILED" : "host_check ok");
    return failures ? 1 : 0;
}
