// Text result writers and matrix printers with the output format of
// CUDA/utils.cuh:32-197.  Host code only: none of these touches a device.
//
// Formats (printf conversions, as in the reference):
//   write_file_cuDoubleComplex  "%f,%f" per entry, ';' between entries, '\n'
//   write_file_double           "%f" per entry, ';' between entries, '\n'
//   write_file_string           the text as is
//   write_matrix_cuDoubleComplex  row by row, "%f,%f," ('f') or "%e,%e,"
//                                ('e') per entry, '\n' per row, then '\n'
//   print_matrix_cuDoubleComplex  stdout, "%.16f %.16fi  " or
//                                "%.16e %.16ei  " per entry
//   print_matrix_double         stdout, "%.4f " or "%.ef " per entry
// The first argument of the writers is the fopen mode ("w" to truncate, "a"
// to append).  The ia_* variants report a file that cannot be opened as
// IA_ERR_IO instead of writing through a NULL stream; the legacy names print
// that error and return.
#include "ia_internal.hpp"

#include <cstdio>

using namespace ia;

namespace {

ia_status open_out(const char *mode, const char *file, FILE **fp) {
    if (!mode || !file) return fail(IA_ERR_ARG, "text writer: NULL mode or file");
    if (mode[0] != 'w' && mode[0] != 'a')
        return fail(IA_ERR_ARG, "text writer: mode must start with 'w' or 'a'");
    *fp = std::fopen(file, mode);
    if (!*fp) return fail(IA_ERR_IO, "text writer: cannot open %s", file);
    return IA_OK;
}

ia_status close_out(FILE *fp, const char *file) {
    const bool bad = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || bad) return fail(IA_ERR_IO, "text writer: write to %s failed", file);
    return IA_OK;
}

void report(const char *what, ia_status st) {
    if (st != IA_OK) std::fprintf(stderr, "%s: %s\n", what, ia_last_error());
}

}  // namespace

extern "C" ia_status ia_write_complex(const char *mode, const char *file, const ia_z *v, int n) {
    if (n < 0 || (n > 0 && !v)) return fail(IA_ERR_ARG, "ia_write_complex: bad arguments");
    FILE *fp = nullptr;
    IA_TRY(open_out(mode, file, &fp));
    for (int i = 0; i < n; ++i) {
        std::fprintf(fp, "%f,%f", v[i].re, v[i].im);
        if (i != n - 1) std::fputc(';', fp);
    }
    std::fputc('\n', fp);
    return close_out(fp, file);
}

extern "C" ia_status ia_write_double(const char *mode, const char *file, const double *v, int n) {
    if (n < 0 || (n > 0 && !v)) return fail(IA_ERR_ARG, "ia_write_double: bad arguments");
    FILE *fp = nullptr;
    IA_TRY(open_out(mode, file, &fp));
    for (int i = 0; i < n; ++i) {
        std::fprintf(fp, "%f", v[i]);
        if (i != n - 1) std::fputc(';', fp);
    }
    std::fputc('\n', fp);
    return close_out(fp, file);
}

extern "C" ia_status ia_write_string(const char *mode, const char *file, const char *text) {
    if (!text) return fail(IA_ERR_ARG, "ia_write_string: NULL text");
    FILE *fp = nullptr;
    IA_TRY(open_out(mode, file, &fp));
    std::fputs(text, fp);
    return close_out(fp, file);
}

extern "C" ia_status ia_write_matrix(const char *fmt, const char *file, const ia_z *a, int r,
                                     int c) {
    if (!fmt || (fmt[0] != 'f' && fmt[0] != 'e'))
        return fail(IA_ERR_ARG, "ia_write_matrix: format must be 'f' or 'e'");
    if (r < 0 || c < 0 || (r > 0 && c > 0 && !a)) return fail(IA_ERR_ARG, "ia_write_matrix: bad sizes");
    FILE *fp = nullptr;
    IA_TRY(open_out("w", file, &fp));
    const char *f = fmt[0] == 'f' ? "%f,%f," : "%e,%e,";
    for (int i = 0; i < r; ++i) {
        for (int j = 0; j < c; ++j) {
            const ia_z z = a[i + (size_t)j * r];
            std::fprintf(fp, f, z.re, z.im);
        }
        std::fputc('\n', fp);
    }
    std::fputc('\n', fp);
    return close_out(fp, file);
}

// ----------------------------------------------------------- legacy names

extern "C" void write_file_cuDoubleComplex(const char *wa, char *file, ia_z *v, int n) {
    report("write_file_cuDoubleComplex", ia_write_complex(wa, file, v, n));
}

extern "C" void write_file_string(const char *wa, char *file, char *text) {
    report("write_file_string", ia_write_string(wa, file, text));
}

extern "C" void write_file_double(const char *wa, char *file, double *v, int n) {
    report("write_file_double", ia_write_double(wa, file, v, n));
}

// utils.cuh:116: always writes "matrix.txt" in the working directory.
extern "C" void write_matrix_cuDoubleComplex(const char *d, ia_z *a, int r, int c) {
    report("write_matrix_cuDoubleComplex", ia_write_matrix(d, "matrix.txt", a, r, c));
}

extern "C" void print_matrix_cuDoubleComplex(const char *d, ia_z *a, int r, int c) {
    if (!d || r < 0 || c < 0 || (r > 0 && c > 0 && !a)) return;
    for (int i = 0; i < r; ++i) {
        for (int j = 0; j < c; ++j) {
            const ia_z z = a[i + (size_t)j * r];
            if (*d == 'f')
                std::printf("%.16f %.16fi  ", z.re, z.im);
            else if (*d == 'e')
                std::printf("%.16e %.16ei  ", z.re, z.im);
        }
        std::printf("\n");
    }
    std::printf("\n");
    std::fflush(stdout);
}

extern "C" void print_matrix_double(const char *d, double *a, int r, int c) {
    if (!d || r < 0 || c < 0 || (r > 0 && c > 0 && !a)) return;
    for (int i = 0; i < r; ++i) {
        for (int j = 0; j < c; ++j) {
            const double x = a[i + (size_t)j * r];
            if (*d == 'f')
                std::printf("%.4f ", x);
            else if (*d == 'e')
                std::printf("%.ef ", x);  // the reference's conversion: precision 0, then 'f'
        }
        std::printf("\n");
    }
    std::printf("\n");
    std::fflush(stdout);
}
