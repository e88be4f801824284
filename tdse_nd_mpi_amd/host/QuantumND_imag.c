/*
 * QuantumND_imag -- imaginary-time relaxation of N_STATES eigenstates on one
 * MI355X.  Drop-in for the reference executable of the same name
 * (src/QuantumND_imag.c:54-263): same command line
 *     QuantumND_imag <params.imag> [KEY VALUE]...
 * same parameter file, same printed summary and the same output.imag file.
 * The propagation, Gram-Schmidt, energies and Lz run in libqnd.
 */
#include "qnd_host.h"
#include "../../include/qnd.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static void check(qnd_ctx *ctx, int rc)
{
    if (rc != 0) {
        fprintf(stderr, "%s\n", qnd_last_error(ctx));
        qnd_abort();
    }
}

static int known_potential(const char *name)
{
    static const char *names[] = {"H_1D", "H_3D", "H_2D", "Ne_2D", "Ar_2D", "Kr_2D", "Xe_2D", "Ne_short_2D", NULL};
    for (int k = 0; names[k]; k++)
        if (strcmp(name, names[k]) == 0) return 1;
    return 0;
}

int main(int argc, char *argv[])
{
    if (argc < 2) {
        fprintf(stderr, "Usage: %s <param_file> [KEY VALUE ...]\n", argv[0]);
        qnd_abort();
    }
    printf("\n##############################################\n");
    printf("\n##    PROGRAM IMAGINARY-TIME PROPAGATION    ##\n");
    printf("\n##############################################\n");
    printf("\n# Parameters read and broadcasted to all ranks\n\n");

    Parameters P;
    memset(&P, 0, sizeof P);
    Initialize_parameters_imag(argv[1], &P, argc, argv);

    double as[20], as_scaled[20];
    if (integrator_coefficients(P.integrator, P.dt, as, as_scaled) < 0) {
        printf("\nUnknown integrator: %s\n", P.integrator);
        perror("Error with the integrator name");
        qnd_abort();
    }
    if (!known_potential(P.potential)) {
        printf("\nUnknown potential: %s\n", P.potential);
        perror("Error with the potential name");
        qnd_abort();
    }
    const int dim = P.dim;
    int64_t N = 1;
    for (int j = 0; j < dim; j++) N *= P.Nr_global[j];

    qnd_ctx *ctx = NULL;
    check(NULL, qnd_create(&ctx, dim, P.Nr_global, P.dr, P.precision, P.N_states));
    check(ctx, qnd_set_potential(ctx, P.potential, P.softcore, NULL));
    check(ctx, qnd_set_integrator(ctx, P.integrator, P.dt));

    double **Psi = (double **)malloc(P.N_states * sizeof(double *));
    double **Ej = (double **)malloc(P.N_states * sizeof(double *));
    double *gauss = NULL;
    if (!Psi || !Ej) {
        perror("malloc with Psi_local");
        qnd_abort();
    }
    for (int j = 0; j < P.N_states; j++) {
        Psi[j] = (double *)calloc(2 * (size_t)N, sizeof(double));
        Ej[j] = (double *)calloc(2, sizeof(double));
        if (!Psi[j] || !Ej[j]) {
            perror("fftw_alloc_complex with Psi_local");
            qnd_abort();
        }
    }
    if (P.guess_mode == 1) {
        gauss = (double *)malloc(2 * (size_t)N * P.N_states * sizeof(double));
        gaussian_stream(P.seed, 2 * N * P.N_states, gauss);
    }
    /* steps of the reference loop for(time=0.;time<T;time+=dt) (Q1) */
    int64_t nsteps = 0;
    for (double t = 0.; t < P.Tintegration; t += P.dt) nsteps++;

    printf("\n# Imaginary-time propagation\n");
    for (int j = 0; j < P.N_states; j++) {
        clock_t start = clock();
        double t0 = wall_seconds();
        /* initial state (src/QuantumND_imag.c:174-179, or the seeded extension) */
        for (int64_t i = 0; i < N; i++) {
            int64_t rem = i, ik[8];
            double rr = 0.;
            for (int k = dim - 1; k >= 0; k--) {
                ik[k] = rem % P.Nr_global[k];
                rem /= P.Nr_global[k];
            }
            for (int k = 0; k < dim; k++) {  /* axis order, as the reference sums */
                double x = -.5 * (P.dr[k] * P.Nr_global[k]) + ik[k] * P.dr[k];
                rr += x * x;
            }
            if (P.guess_mode == 0) {
                double radius = sqrt(rr);
                Psi[j][2 * i] = (radius + (double)(i % 3)) * exp(-radius);
                Psi[j][2 * i + 1] = 0.;
            } else {
                const double *g = gauss + 2 * ((size_t)j * N + i);
                const double e = exp(-rr / 20.);
                Psi[j][2 * i] = g[0] * e;
                Psi[j][2 * i + 1] = g[1] * e;
            }
        }
        check(ctx, qnd_upload_state(ctx, j, Psi[j]));
        check(ctx, qnd_imag_steps(ctx, j, nsteps, 1));
        check(ctx, qnd_energy_moments(ctx, j, 2, Ej[j]));
        double t1 = wall_seconds();
        int code = QND_OBS_LZ;
        double Lz = 0.;
        check(ctx, qnd_observables(ctx, j, &code, 1, NULL, NULL, &Lz));
        printf("\n> eigenfunction %d:\n", j + 1);
        printf("  E\t\t\t%16.15f\n", Ej[j][0]);
        printf("  DE\t\t\t%5.4e \n", Ej[j][1] - Ej[j][0] * Ej[j][0]);
        printf("  Lz\t\t\t%5.4e \n", Lz);
        printf("  Elapsed CPU time\t%16.15f seconds\n", (double)(clock() - start) / CLOCKS_PER_SEC);
        printf("  Elapsed real time\t%16.15f seconds\n", t1 - t0);
        fflush(stdout);
    }
    printf("\n");

    /* global phase at the first strict maximum of |psi| (src/QuantumND_imag.c:215-240) */
    for (int j = 0; j < P.N_states; j++) {
        check(ctx, qnd_download_state(ctx, j, Psi[j]));
        double rho = 0., phase = 0.;
        for (int64_t i = 0; i < N; i++) {
            double complex z = Psi[j][2 * i] + I * Psi[j][2 * i + 1];
            double m = cabs(z);
            if (m > rho) {
                rho = m;
                phase = carg(z);
            }
        }
        double complex fac = cexp(-I * phase);
        for (int64_t i = 0; i < N; i++) {
            double complex z = (Psi[j][2 * i] + I * Psi[j][2 * i + 1]) * fac;
            Psi[j][2 * i] = creal(z);
            Psi[j][2 * i + 1] = cimag(z);
        }
    }
    Write_output_imag(P.output_imag, Psi, Ej, &P, N);
    printf("# The code ended successfully\n\n");

    qnd_destroy(ctx);
    for (int j = 0; j < P.N_states; j++) {
        free(Psi[j]);
        free(Ej[j]);
    }
    free(Psi);
    free(Ej);
    free(gauss);
    Finalize_parameters(&P);
    return 0;
}
