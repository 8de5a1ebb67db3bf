/*
 * matload.c -- reading a sparse matrix stored with PETSc's binary viewer
 * (the MatLoad format of the reference's share/slepc/datafiles/matrices/ files,
 * read by src/eps/tutorials/ex4.c:43-50 through PetscViewerBinaryOpen + MatLoad).
 *
 * The format (PETSc MatView_SeqAIJ_Binary / MatLoad_SeqAIJ_Binary): all
 * integers 32-bit and all values IEEE doubles, big-endian;
 *   int32  classid = 1211216 (MAT_FILE_CLASSID)
 *   int32  M, N, nz          (global rows, columns, stored entries)
 *   int32  rowlen[M]         (entries per row)
 *   int32  colidx[nz]        (column indices, row by row)
 *   double val[nz]
 * Rows are returned in CSR with their column indices sorted (stable for
 * duplicates, which are kept as stored).
 */
#include "slepcimpl.h"

#define MAT_FILE_CLASSID 1211216

static uint32_t be32(const unsigned char *p)
{
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

static double be64d(const unsigned char *p)
{
  uint64_t u = 0;
  double d;
  for (int i = 0; i < 8; i++) u = (u << 8) | p[i];
  memcpy(&d, &u, sizeof(d));
  return d;
}

/* insertion sort of one row's (col, val) pairs: rows are short */
static void SortRow(PetscInt *c, PetscScalar *v, PetscInt len)
{
  for (PetscInt i = 1; i < len; i++) {
    PetscInt ci = c[i];
    PetscScalar vi = v[i];
    PetscInt j = i - 1;
    while (j >= 0 && c[j] > ci) {
      c[j + 1] = c[j];
      v[j + 1] = v[j];
      j--;
    }
    c[j + 1] = ci;
    v[j + 1] = vi;
  }
}

PetscErrorCode MatReadPETScBinary(const char *path, PetscInt *m, PetscInt *n, PetscInt **rowptr, PetscInt **colidx,
                                  PetscScalar **val)
{
  FILE *f;
  unsigned char hdr[16], *buf = NULL;
  uint32_t classid, M, N, nz;
  size_t need, got;
  PetscInt *rp = NULL, *ci = NULL;
  PetscScalar *va = NULL;
  PetscCheck(path && m && n && rowptr && colidx && val, PETSC_ERR_ARG_NULL, "null argument");
  f = fopen(path, "rb");
  PetscCheck(f, PETSC_ERR_FILE_OPEN, "cannot open %s", path);
  got = fread(hdr, 1, 16, f);
  if (got != 16) { fclose(f); SETERRQ(PETSC_ERR_FILE_READ, "%s: short header", path); }
  classid = be32(hdr);
  M = be32(hdr + 4);
  N = be32(hdr + 8);
  nz = be32(hdr + 12);
  if (classid != MAT_FILE_CLASSID) { fclose(f); SETERRQ(PETSC_ERR_FILE_UNEXPECTED, "%s: not a PETSc binary matrix (class id %u)", path, classid); }
  if (M == 0 || N == 0 || M > 0x7fffffffu || N > 0x7fffffffu || nz > 0x7fffffffu) {
    fclose(f);
    SETERRQ(PETSC_ERR_FILE_UNEXPECTED, "%s: bad sizes %u x %u, %u entries", path, M, N, nz);
  }
  need = 4 * (size_t)M + 12 * (size_t)nz;
  buf = (unsigned char *)malloc(need ? need : 1);
  if (!buf) { fclose(f); SETERRQ(PETSC_ERR_MEM, "out of host memory"); }
  got = fread(buf, 1, need, f);
  fclose(f);
  if (got != need) { free(buf); SETERRQ(PETSC_ERR_FILE_READ, "%s: truncated (%zu of %zu bytes of data)", path, got, need); }
  rp = (PetscInt *)malloc(sizeof(PetscInt) * ((size_t)M + 1));
  ci = (PetscInt *)malloc(sizeof(PetscInt) * (nz ? nz : 1));
  va = (PetscScalar *)malloc(sizeof(PetscScalar) * (nz ? nz : 1));
  if (!rp || !ci || !va) { free(buf); free(rp); free(ci); free(va); SETERRQ(PETSC_ERR_MEM, "out of host memory"); }
  rp[0] = 0;
  for (uint32_t i = 0; i < M; i++) {
    uint32_t len = be32(buf + 4 * (size_t)i);
    if ((uint64_t)rp[i] + len > nz) { free(buf); free(rp); free(ci); free(va); SETERRQ(PETSC_ERR_FILE_UNEXPECTED, "%s: row lengths exceed %u entries", path, nz); }
    rp[i + 1] = rp[i] + (PetscInt)len;
  }
  if ((uint32_t)rp[M] != nz) { free(buf); free(rp); free(ci); free(va); SETERRQ(PETSC_ERR_FILE_UNEXPECTED, "%s: row lengths sum to %d, header says %u", path, rp[M], nz); }
  for (uint32_t p = 0; p < nz; p++) {
    uint32_t c = be32(buf + 4 * (size_t)M + 4 * (size_t)p);
    if (c >= N) { free(buf); free(rp); free(ci); free(va); SETERRQ(PETSC_ERR_FILE_UNEXPECTED, "%s: column %u out of range", path, c); }
    ci[p] = (PetscInt)c;
    va[p] = be64d(buf + 4 * (size_t)M + 4 * (size_t)nz + 8 * (size_t)p);
  }
  free(buf);
  for (uint32_t i = 0; i < M; i++) SortRow(ci + rp[i], va + rp[i], rp[i + 1] - rp[i]);
  *m = (PetscInt)M;
  *n = (PetscInt)N;
  *rowptr = rp;
  *colidx = ci;
  *val = va;
  return PETSC_SUCCESS;
}

void MatFreePETScBinary(PetscInt *rowptr, PetscInt *colidx, PetscScalar *val)
{
  free(rowptr);
  free(colidx);
  free(val);
}

/* MatLoad of a square matrix into a MATHIPCSR (the CSR goes to the device) */
PetscErrorCode MatLoadPETScBinary(const char *path, Mat *A)
{
  PetscInt m, n, *rp, *ci;
  PetscScalar *va;
  PetscErrorCode ierr;
  PetscCall(MatReadPETScBinary(path, &m, &n, &rp, &ci, &va));
  if (m != n) {
    MatFreePETScBinary(rp, ci, va);
    SETERRQ(PETSC_ERR_SUP, "%s: %d x %d matrix; the eigensolver takes square matrices", path, m, n);
  }
  ierr = MatCreateHIPCSR(m, rp, ci, va, A);
  MatFreePETScBinary(rp, ci, va);
  PetscCall(ierr);
  return PETSC_SUCCESS;
}
