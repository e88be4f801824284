/*
 * qnd_io.c -- file formats of the QuantumND drop-in, byte-compatible with
 * src/io_functions.c for a single process: output.imag (text header + FP64
 * complex binary), observables.real, wavefunction.real, PED.real, PMD.real,
 * and the centred read-back of output.imag.  The soft-core extension travels
 * in a side file (<output.imag>.softcore), so the reference can read every
 * output.imag this code writes.
 */
#include "qnd_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Name of the side file that carries the soft-core extension next to an
 * output.imag file: <output.imag>.softcore.  The header itself keeps the
 * reference's keys only, because the reference reader stops at an unknown
 * key (src/io_functions.c:377-385) and then cannot find the PSI marker. */
void softcore_side_file(const char *output_imag, char *buf, size_t len)
{
    snprintf(buf, len, "%s.softcore", output_imag);
}

/* src/io_functions.c:57-106 */
void Write_output_imag(const char *filename, double **Psi, double **Ej, const Parameters *P, int64_t N)
{
    char side[MAX_OUTPUT_FILES_CHAR + 16];
    softcore_side_file(filename, side, sizeof side);
    if (P->softcore >= 0) {
        FILE *s = fopen(side, "w");
        if (!s) {
            perror("Error opening the soft-core side file");
            qnd_abort();
        }
        fprintf(s, "SOFTCORE\t%16.15f\n", P->softcore);
        fclose(s);
    } else {
        remove(side);  /* a stale side file from an earlier run must not apply */
    }
    FILE *f = fopen(filename, "w");
    if (!f) {
        perror("Error opening the file in write_output_imag");
        qnd_abort();
    }
    fprintf(f, "DIMENSION\t%d\n", P->dim);
    fprintf(f, "N_STATES\t%d\n", P->N_states);
    fprintf(f, "NR");
    for (int j = 0; j < P->dim; j++) fprintf(f, "\t%lld", (long long)P->Nr_global[j]);
    fprintf(f, "\nDR");
    for (int j = 0; j < P->dim; j++) fprintf(f, "\t%16.15f", P->dr[j]);
    fprintf(f, "\nPOTENTIAL\t%s\n", P->potential);
    fprintf(f, "E\t");
    for (int j = 0; j < P->N_states; j++) fprintf(f, "\t%16.15f", Ej[j][0]);
    fprintf(f, "\nDE");
    for (int j = 0; j < P->N_states; j++) fprintf(f, "\t%16.15f", Ej[j][1] - Ej[j][0] * Ej[j][0]);
    fprintf(f, "\nPSI\n");
    for (int j = 0; j < P->N_states; j++) {
        size_t w = fwrite(Psi[j], 2 * sizeof(double), (size_t)N, f);
        if (w != (size_t)N) {
            fprintf(stderr, "Rank 0 ERROR: Failed to write all %lld elements to binary file.\n", (long long)N);
            qnd_abort();
        }
    }
    fclose(f);
}

/* src/io_functions.c:109-133 */
void Write_observables(FILE *f, double time, const double *values, int nvalues)
{
    if (!f) return;
    fprintf(f, "%16.15f", time);
    for (int k = 0; k < nvalues; k++) fprintf(f, "\t%16.15f", values[k]);
    fprintf(f, "\n");
    fflush(f);
}

/* src/io_functions.c:13-54 */
void Write_wavefunction_to_file(const char *filename, const double *Psi, const Parameters *P, int64_t N)
{
    FILE *f = fopen(filename, "w");
    if (!f) {
        perror("Error opening the file in write_wavefunction_imag_to_file");
        qnd_abort();
    }
    fprintf(f, "DIMENSION\t%d\nNR", P->dim);
    for (int j = 0; j < P->dim; j++) fprintf(f, "\t%lld", (long long)P->Nr_global[j]);
    fprintf(f, "\nDR");
    for (int j = 0; j < P->dim; j++) fprintf(f, "\t%16.15f", P->dr[j]);
    fprintf(f, "\nPSI_BINARY_START\n");
    size_t w = fwrite(Psi, 2 * sizeof(double), (size_t)N, f);
    if (w != (size_t)N) {
        fprintf(stderr, "Rank 0 failed to write all elements. Expected %lld, wrote %zu.\n", (long long)N, w);
        perror("Error writing Psi array to binary file");
        qnd_abort();
    }
    fclose(f);
}

/* src/io_functions.c:201-209 */
void Write_PED(const double *S, int Ne, double de)
{
    FILE *f = fopen("PED.real", "w");
    if (!f) {
        perror("Error opening the file in write_output_imag");
        qnd_abort();
    }
    for (int j = 0; j < Ne; j++) fprintf(f, "%16.15f\t%16.15f\n", j * de, S[j]);
    fclose(f);
}

static double axis_p(int64_t i, int64_t n, double dr)
{
    const double dp = 2. * M_PI / (dr * n);
    return (i < n / 2) ? i * dp : (i - n) * dp;
}

/* src/io_functions.c:217-286: header, then (|psi|^2, arg psi) for every
 * point whose momenta all satisfy |p_j| <= P_MAX_j, in native FFT order. */
void Write_PMD(const double *psi_p, const Parameters *P)
{
    const int dim = P->dim;
    int64_t N = 1, cnt[8];
    for (int j = 0; j < dim; j++) {
        N *= P->Nr_global[j];
        cnt[j] = 0;
        for (int64_t i = 0; i < P->Nr_global[j]; i++)
            if (fabs(axis_p(i, P->Nr_global[j], P->dr[j])) <= P->pr_max[j]) cnt[j]++;
    }
    FILE *f = fopen("PMD.real", "w");
    if (!f) {
        perror("Error opening the PMD.real file");
        qnd_abort();
    }
    fprintf(f, "DIMENSION\t%d\nNPR", dim);
    for (int j = 0; j < dim; j++) fprintf(f, "\t%lld", (long long)cnt[j]);
    fprintf(f, "\nDP");
    for (int j = 0; j < dim; j++) fprintf(f, "\t%16.15f", 2. * M_PI / (double)(P->dr[j] * P->Nr_global[j]));
    fprintf(f, "\nDATA_START\n");
    fclose(f);
    double *buf = (double *)malloc(2 * (size_t)N * sizeof(double));
    if (!buf) {
        fprintf(stderr, "Rank 0 ERROR: Memory allocation failed for PMD buffer.\n");
        qnd_abort();
    }
    size_t b = 0;
    for (int64_t i = 0; i < N; i++) {
        int64_t rem = i;
        int inside = 1;
        for (int j = dim - 1; j >= 0; j--) {
            int64_t ij = rem % P->Nr_global[j];
            rem /= P->Nr_global[j];
            if (fabs(axis_p(ij, P->Nr_global[j], P->dr[j])) > P->pr_max[j]) inside = 0;
        }
        if (inside) {
            const double re = psi_p[2 * i], im = psi_p[2 * i + 1];
            const double m = hypot(re, im);
            buf[b++] = m * m;
            buf[b++] = atan2(im, re);
        }
    }
    f = fopen("PMD.real", "ab");
    if (!f) {
        perror("Error opening appended PMD.bin file");
        qnd_abort();
    }
    if (fwrite(buf, sizeof(double), b, f) != b) fprintf(stderr, "Rank 0 ERROR: Failed to write complete PMD record.\n");
    fclose(f);
    free(buf);
}

/* src/io_functions.c:292-459: locate the binary block after the PSI marker,
 * embed each requested imaginary-time state centred in the real-time grid
 * (offset (NR_real - NR_imag)/2 per axis, points outside dropped) and
 * accumulate coeff_k * psi_k. */
void Read_wavefunction(const Parameters *P, const Parameters *P_imag, double *Psi)
{
    char key[32];
    const int dim = P->dim;
    long psi_offset = 0;
    FILE *f = fopen(P->output_imag, "rb");
    if (!f) {
        perror("Error opening file for header read");
        qnd_abort();
    }
    while (fscanf(f, "%31s", key) != EOF) {
        if (strcmp(key, "PSI") == 0) {
            int ch;
            while ((ch = fgetc(f)) != '\n' && ch != EOF) {
            }
            psi_offset = ftell(f);
            break;
        }
        if (strcmp(key, "DIMENSION") == 0 || strcmp(key, "N_STATES") == 0) {
            int v;
            if (fscanf(f, "%d", &v) != 1) break;
        } else if (strcmp(key, "POTENTIAL") == 0) {
            if (fscanf(f, "%31s", key) != 1) break;
        } else if (strcmp(key, "NR") == 0) {
            long long v;
            for (int j = 0; j < P_imag->dim; j++)
                if (fscanf(f, "%lld", &v) != 1) break;
        } else if (strcmp(key, "DR") == 0) {
            double v;
            for (int j = 0; j < P_imag->dim; j++)
                if (fscanf(f, "%lf", &v) != 1) break;
        } else if (strcmp(key, "E") == 0 || strcmp(key, "DE") == 0) {
            double v;
            for (int j = 0; j < P_imag->N_states; j++)
                if (fscanf(f, "%lf", &v) != 1) break;
        } else {
            printf("Warning: Skipping unknown key %s in Read_wavefunction_MPI\n", key);
            if (fscanf(f, "%*s") != 0) break;
        }
    }
    if (psi_offset == 0) {
        fclose(f);
        fprintf(stderr, "Rank 0: Error: Could not find 'PSI' marker in file.\n");
        qnd_abort();
    }
    int64_t Nimag = 1, Nreal = 1;
    int64_t off[8];
    for (int j = 0; j < dim; j++) {
        Nimag *= P_imag->Nr_global[j];
        Nreal *= P->Nr_global[j];
        off[j] = (P->Nr_global[j] - P_imag->Nr_global[j]) / 2;
    }
    double *buf = (double *)malloc(2 * (size_t)Nimag * sizeof(double));
    if (!buf) {
        fprintf(stderr, "Rank 0: Memory allocation error for temp_psi_local\n");
        qnd_abort();
    }
    for (int k = 0; k < P->N_states; k++) {
        const long pos = psi_offset + (long)(P->j_states[k] - 1) * (long)Nimag * 16L;
        if (fseek(f, pos, SEEK_SET) != 0 || fread(buf, 16, (size_t)Nimag, f) != (size_t)Nimag) {
            fprintf(stderr, "Rank 0: MPI Read Error during chunking\n");
            qnd_abort();
        }
        const double cr = P->coeffs_states[2 * k], ci = P->coeffs_states[2 * k + 1];
        for (int64_t i = 0; i < Nimag; i++) {
            int64_t rem = i, ind = 0, mult = 1;
            int valid = 1;
            for (int j = dim - 1; j >= 0; j--) {
                const int64_t ij = rem % P_imag->Nr_global[j];
                rem /= P_imag->Nr_global[j];
                const int64_t g = ij + off[j];
                if (g < 0 || g >= P->Nr_global[j]) {
                    valid = 0;
                    break;
                }
                ind += g * mult;
                mult *= P->Nr_global[j];
            }
            if (valid) {
                const double re = buf[2 * i], im = buf[2 * i + 1];
                Psi[2 * ind] += cr * re - ci * im;
                Psi[2 * ind + 1] += cr * im + ci * re;
            }
        }
    }
    (void)Nreal;
    free(buf);
    fclose(f);
}
