/*
 * ref_tridiag_driver.c -- runs the reference's own pll_tridiag (compiled
 * from pll_tridiag.c where it lies, see Makefile target `ref`) on systems
 * read from a file, at one MPI rank.  TEST INFRASTRUCTURE ONLY.
 *
 * input : int32 nsys, int32 n, then a, b, c, r as nsys*n doubles each
 *         (line-major: system s occupies [s*n, (s+1)*n))
 * output: b, c, r after pll_tridiag (same layout), i.e. the eliminated
 *         factors and the solution (pll_tridiag.c:21-47, 96-110)
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
    double *acof, *bcof, *ccof, *rhs;
} Tds;

int Rank, Size;   /* the extern globals pll_tridiag.c reads (pll_tridiag.c:9-10) */
void pll_tridiag(Tds *arr, int N, int n);

int main(int argc, char **argv)
{
    int nsys, n, s;
    FILE *f;
    double *buf;
    Tds *arr;
    if (argc < 3) { fprintf(stderr, "usage: ref_tridiag in.bin out.bin\n"); return 2; }
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &Size);
    f = fopen(argv[1], "rb");
    if (!f || fread(&nsys, 4, 1, f) != 1 || fread(&n, 4, 1, f) != 1) { fprintf(stderr, "bad input\n"); return 1; }
    buf = (double *)malloc(sizeof(double) * 4 * (size_t)nsys * n);
    if (fread(buf, sizeof(double), 4 * (size_t)nsys * n, f) != 4 * (size_t)nsys * n) { fprintf(stderr, "short input\n"); return 1; }
    fclose(f);
    arr = (Tds *)malloc(sizeof(Tds) * nsys);
    for (s = 0; s < nsys; s++) {
        arr[s].acof = buf + 0 * (size_t)nsys * n + (size_t)s * n;
        arr[s].bcof = buf + 1 * (size_t)nsys * n + (size_t)s * n;
        arr[s].ccof = buf + 2 * (size_t)nsys * n + (size_t)s * n;
        arr[s].rhs = buf + 3 * (size_t)nsys * n + (size_t)s * n;
    }
    pll_tridiag(arr, nsys, n);
    f = fopen(argv[2], "wb");
    fwrite(buf + (size_t)nsys * n, sizeof(double), 3 * (size_t)nsys * n, f);
    fclose(f);
    MPI_Finalize();
    return 0;
}
