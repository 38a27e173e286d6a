// PETSc binary matrix dumps: the file DumpMatrixEquation writes
// (femtools/Petsc_Tools.F90:1487-1504: MatView(A), VecView(b), VecView(x0)
// on a binary viewer) when a solve fails (femtools/Solvers.F90:1369-1479),
// and that tools/petsc_readnsolve.F90:60-83 reads back with MatLoad /
// VecLoad.  PETSc's published binary layout, 32-bit indices (Fluidity's
// PETSc builds), big-endian throughout:
//   matrix: int32 MAT_FILE_CLASSID (1211216), rows, cols, nnz;
//           int32 row lengths[rows]; int32 column indices[nnz] (0-based,
//           ascending within a row); float64 values[nnz]
//   vector: int32 VEC_FILE_CLASSID (1211214), n; float64 values[n]
// Host code only: the arrays it fills are what flgpu_sparsity_bind and
// flgpu_matrix_set take, so a dump replays through the device solver.
#include <cstdio>
#include <cstring>
#include <vector>

#include "flgpu_internal.h"

namespace flgpu {
namespace {

constexpr int32_t kMatClassId = 1211216, kVecClassId = 1211214;

uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

uint64_t bswap64(uint64_t v) {
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

bool host_little_endian() {
    const uint16_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

struct Reader {
    FILE* f = nullptr;
    bool ok = true;
    ~Reader() {
        if (f) std::fclose(f);
    }
    bool ints(int32_t* out, size_t n) {
        if (n && std::fread(out, sizeof(int32_t), n, f) != n) return ok = false;
        if (host_little_endian())
            for (size_t k = 0; k < n; k++) {
                uint32_t u;
                std::memcpy(&u, &out[k], 4);
                u = bswap32(u);
                std::memcpy(&out[k], &u, 4);
            }
        return true;
    }
    bool reals(double* out, size_t n) {
        if (n && std::fread(out, sizeof(double), n, f) != n) return ok = false;
        if (host_little_endian())
            for (size_t k = 0; k < n; k++) {
                uint64_t u;
                std::memcpy(&u, &out[k], 8);
                u = bswap64(u);
                std::memcpy(&out[k], &u, 8);
            }
        return true;
    }
};

struct Writer {
    FILE* f = nullptr;
    ~Writer() {
        if (f) std::fclose(f);
    }
    bool ints(const int32_t* in, size_t n) {
        std::vector<uint32_t> b(n);
        for (size_t k = 0; k < n; k++) {
            uint32_t u;
            std::memcpy(&u, &in[k], 4);
            b[k] = host_little_endian() ? bswap32(u) : u;
        }
        return !n || std::fwrite(b.data(), 4, n, f) == n;
    }
    bool reals(const double* in, size_t n) {
        std::vector<uint64_t> b(n);
        for (size_t k = 0; k < n; k++) {
            uint64_t u;
            std::memcpy(&u, &in[k], 8);
            b[k] = host_little_endian() ? bswap64(u) : u;
        }
        return !n || std::fwrite(b.data(), 8, n, f) == n;
    }
};

}  // namespace
}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_matrixdump_read(const char* filename, int32_t* n, int64_t* nnz, int32_t* findrm, int32_t* colm,
                                     double* val, double* b, double* x0, int32_t* has_x0) {
    return guarded("flgpu_matrixdump_read", [&]() -> int {
        if (!filename || !n || !nnz) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: null argument");
        Reader r;
        r.f = std::fopen(filename, "rb");
        if (!r.f) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: cannot open %s", filename);
        int32_t h[4];
        if (!r.ints(h, 4) || h[0] != kMatClassId)
            return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: %s does not start with a PETSc binary matrix",
                             filename);
        const int32_t rows = h[1], cols = h[2], total = h[3];
        if (rows <= 0 || rows != cols || total < 0)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu_matrixdump_read: %d x %d matrix with %d entries (square only)",
                             rows, cols, total);
        *n = rows;
        *nnz = total;
        if (!findrm) return FLGPU_OK;  // size query
        if (!colm || !val || !b) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: null output array");
        std::vector<int32_t> len(rows);
        if (!r.ints(len.data(), rows)) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: truncated row lengths");
        findrm[0] = 1;
        int64_t sum = 0;  // 64-bit, so a corrupt length cannot wrap the running offset
        for (int i = 0; i < rows; i++) {
            if (len[i] < 0) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: negative row length");
            sum += len[i];
            if (sum > total)
                return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: row lengths exceed the %d entries of the header",
                                 total);
            findrm[i + 1] = (int32_t)(sum + 1);
        }
        if (sum != total)
            return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: row lengths sum to %lld, header says %d",
                             (long long)sum, total);
        if (!r.ints(colm, (size_t)total))
            return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: truncated column indices");
        for (int i = 0; i < rows; i++)
            for (int k = findrm[i] - 1; k < findrm[i + 1] - 1; k++)
                if (colm[k] < 0 || colm[k] >= cols || (k > findrm[i] - 1 && colm[k] <= colm[k - 1]))
                    return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: row %d: columns out of range or unsorted",
                                     i + 1);
        for (int32_t k = 0; k < total; k++) colm[k] += 1;  // 1-based, as csr_sparsity%colm
        if (!r.reals(val, (size_t)total)) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: truncated values");
        int32_t vh[2];
        if (!r.ints(vh, 2) || vh[0] != kVecClassId || vh[1] != rows)
            return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: right-hand side missing or of the wrong size");
        if (!r.reals(b, rows)) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: truncated right-hand side");
        int32_t got_x0 = 0;
        if (r.ints(vh, 2)) {  // the initial guess is optional (older dumps stop after b)
            if (vh[0] != kVecClassId || vh[1] != rows)
                return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: initial guess of the wrong size");
            std::vector<double> tmp(rows);
            if (!r.reals(x0 ? x0 : tmp.data(), rows))
                return set_error(FLGPU_EINVAL, "flgpu_matrixdump_read: truncated initial guess");
            got_x0 = 1;
        }
        if (has_x0) *has_x0 = got_x0;
        return FLGPU_OK;
    });
}

extern "C" int flgpu_matrixdump_write(const char* filename, int32_t n, const int32_t* findrm, const int32_t* colm,
                                      const double* val, const double* b, const double* x0) {
    return guarded("flgpu_matrixdump_write", [&]() -> int {
        if (!filename || !findrm || !colm || !val || !b || n <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: null argument");
        if (findrm[0] != 1) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: findrm(1) != 1");
        const int32_t total = findrm[n] - 1;
        std::vector<int32_t> len(n), cols0((size_t)total);
        for (int i = 0; i < n; i++) {
            len[i] = findrm[i + 1] - findrm[i];
            if (len[i] < 0) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: findrm decreasing at %d", i + 1);
        }
        for (int32_t k = 0; k < total; k++) {
            if (colm[k] < 1 || colm[k] > n) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: colm(%d)", k + 1);
            cols0[k] = colm[k] - 1;
        }
        Writer w;
        w.f = std::fopen(filename, "wb");
        if (!w.f) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: cannot create %s", filename);
        const int32_t mh[4] = {kMatClassId, n, n, total};
        const int32_t vh[2] = {kVecClassId, n};
        bool ok = w.ints(mh, 4) && w.ints(len.data(), n) && w.ints(cols0.data(), (size_t)total) &&
                  w.reals(val, (size_t)total) && w.ints(vh, 2) && w.reals(b, n);
        if (ok && x0) ok = w.ints(vh, 2) && w.reals(x0, n);
        if (!ok) return set_error(FLGPU_EINVAL, "flgpu_matrixdump_write: write to %s failed", filename);
        return FLGPU_OK;
    });
}
