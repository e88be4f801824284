/*
 * QuantumND_real -- real-time propagation on one MI355X.  Drop-in for the
 * reference executable of the same name (src/QuantumND_real.c:47-264): same
 * command line
 *     QuantumND_real <params.real> [KEY VALUE]...
 * same parameter file and output.imag input, and the same output files
 * (electric_field.real, vector_potential.real, observables.real, PED.real,
 * PMD.real, wavefunction.real).  Propagation and observables run in libqnd;
 * the observables are recorded every step, as the reference does.
 */
#include "qnd_host.h"
#include "../../include/qnd.h"

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
    printf("\n##      PROGRAM REAL-TIME PROPAGATION       ##\n");
    printf("\n##############################################\n");
    printf("\n# Parameters read and broadcasted to all ranks\n\n");

    Parameters P, P_imag;
    Field F;
    memset(&P, 0, sizeof P);
    memset(&P_imag, 0, sizeof P_imag);
    memset(&F, 0, sizeof F);
    Initialize_parameters_real(argv[1], &P, &P_imag, &F, argc, argv);

    double as[20], as_scaled[20];
    const int s = integrator_coefficients(P.integrator, P.dt, as, as_scaled);
    if (s < 0) {
        printf("\nUnknown integrator: %s\n", P.integrator);
        perror("Error with the integrator name");
        qnd_abort();
    }
    Initialize_field_vectors(&P, &F, as_scaled, s);

    /* observables (src/calculate_observables.c:12-83) */
    const int dim = P.dim;
    int nobs = P.N_observables, width = 0;
    int *codes = (int *)calloc(nobs > 0 ? nobs : 1, sizeof(int));
    int *dims = (int *)calloc(nobs > 0 ? nobs : 1, sizeof(int));
    int want_pmd = 0, want_ped = 0;
    for (int j = 0; j < nobs; j++) {
        const char *nm = P.observables[j];
        if (!strcmp(nm, "PMD") || !strcmp(nm, "PED") || !strcmp(nm, "wavefunction") || !strcmp(nm, "none")) {
            dims[j] = 0;
        } else if (observable_code(nm) == 0) {
            printf("Unknown observable: %s. Skipping.\n", nm);
            dims[j] = 0;
        } else {
            dims[j] = observable_width(nm, dim);
        }
        codes[j] = dims[j] ? observable_code(nm) : 0;
        width += dims[j];
    }
    FILE *fobs = fopen("observables.real", "w");
    if (!fobs) {
        perror("Error opening the file of observables");
        qnd_abort();
    }
    /* header as src/QuantumND_real.c:100-102, including its dimensions[k]
     * indexing; an index past the list (undefined in the reference) prints
     * the name. */
    fprintf(fobs, "time");
    for (int j = 0; j < nobs; j++)
        for (int k = 0; k < dims[j]; k++)
            if (k >= nobs || dims[k] != 0) fprintf(fobs, "\t%s", P.observables[j]);
    fprintf(fobs, "\n");

    if (!known_potential(P.potential)) {
        printf("\nUnknown potential: %s\n", P.potential);
        perror("Error with the potential name");
        qnd_abort();
    }
    int64_t N = 1;
    for (int j = 0; j < dim; j++) N *= P.Nr_global[j];

    printf("\n# Initialization of the wave function\n");
    double *Psi = (double *)calloc(2 * (size_t)N, sizeof(double));
    if (!Psi) {
        fprintf(stderr, "ERROR: memory allocation failed for Psi.\n");
        qnd_abort();
    }
    Read_wavefunction(&P, &P_imag, Psi);

    qnd_ctx *ctx = NULL;
    check(NULL, qnd_create(&ctx, dim, P.Nr_global, P.dr, P.precision, 1));
    check(ctx, qnd_set_potential(ctx, P.potential, P.softcore, P.cap));
    check(ctx, qnd_set_integrator(ctx, P.integrator, P.dt));
    check(ctx, qnd_set_mask(ctx, P.mask_frac));
    check(ctx, qnd_upload_state(ctx, 0, Psi));
    {   /* initial-state printout (src/io_functions.c:461-490) */
        int cn = QND_OBS_NORM, cl = QND_OBS_LZ;
        double nrm = 0., E2[2] = {0, 0}, Lz = 0.;
        check(ctx, qnd_observables(ctx, 0, &cn, 1, NULL, NULL, &nrm));
        check(ctx, qnd_energy_moments(ctx, 0, 2, E2));
        check(ctx, qnd_observables(ctx, 0, &cl, 1, NULL, NULL, &Lz));
        printf("\n  norm\t\t\t%16.15f\n  E\t\t\t%16.15f\n  DE\t\t\t%5.4e\n  Lz\t\t\t%16.15f\n\n", sqrt(nrm), E2[0],
               E2[1] - E2[0] * E2[0], Lz);
    }

    /* field at the step times for the observables (src/QuantumND_real.c:173) */
    double *Eobs = (double *)calloc((size_t)(P.N_time > 0 ? P.N_time : 1) * dim, sizeof(double));
    double *Aobs = (double *)calloc((size_t)(P.N_time > 0 ? P.N_time : 1) * dim, sizeof(double));
    for (int64_t i = 0; i < P.N_time; i++) Field_vectors(&P, &F, P.time[i], &Aobs[i * dim], &Eobs[i * dim]);
    check(ctx, qnd_set_field(ctx, P.N_time, F.E, F.A, Eobs, Aobs));

    printf("\n# Real-time propagation\n");
    clock_t start = clock();
    double t0 = wall_seconds();
    int gauge = -1;
    if (strcmp(F.gauge, "velocity") == 0) gauge = QND_GAUGE_VELOCITY;
    else if (strcmp(F.gauge, "length") == 0) gauge = QND_GAUGE_LENGTH;
    int percent = 0;
    if (gauge >= 0) {
        const int64_t chunk = P.chunk;
        double *rec = (double *)calloc((size_t)chunk * (width > 0 ? width : 1), sizeof(double));
        for (int64_t it0 = 0; it0 < P.N_time; it0 += chunk) {
            const int64_t n = (P.N_time - it0) < chunk ? (P.N_time - it0) : chunk;
            check(ctx, qnd_real_steps(ctx, 0, gauge, it0, n, codes, nobs, width > 0 ? rec : NULL));
            for (int64_t k = 0; k < n; k++) {
                const int64_t it = it0 + k;
                if (nobs > 0) Write_observables(fobs, P.time[it], rec + k * width, width);
                if (percent < 100 * it / P.N_time) {
                    printf("> %d%%\n", percent);
                    percent += 10;
                }
            }
        }
        free(rec);
    }
    double t1 = wall_seconds();
    printf("  Elapsed time\t\t%16.15f seconds\n", (double)(clock() - start) / CLOCKS_PER_SEC);
    printf("  Elapsed real time\t%16.15f seconds\n", t1 - t0);
    fclose(fobs);

    /* final observables (src/io_functions.c:136-287) */
    printf("\n# Calculation of the final observables\n");
    for (int j = 0; j < nobs; j++) {
        if (!strcmp(P.observables[j], "PED")) want_ped = 1;
        else if (!strcmp(P.observables[j], "PMD")) want_pmd = 1;
        else if (!strcmp(P.observables[j], "wavefunction")) {
            check(ctx, qnd_download_state(ctx, 0, Psi));
            Write_wavefunction_to_file("wavefunction.real", Psi, &P, N);
        }
    }
    if (want_ped || want_pmd) {
        double *ped = want_ped ? (double *)calloc(P.Ne > 0 ? P.Ne : 1, sizeof(double)) : NULL;
        check(ctx, qnd_final_observables(ctx, 0, P.R_mask, P.pr_max, want_ped ? P.Ne : 0, P.de, ped,
                                         want_pmd ? Psi : NULL));
        if (want_ped) {
            printf("> Calculation of the photoelectron energy distribution\n");
            Write_PED(ped, P.Ne, P.de);
        }
        if (want_pmd) {
            printf("> Calculation of the photoelectron momentum distribution\n");
            Write_PMD(Psi, &P);
        }
        free(ped);
    }
    printf("\n# The code ended successfully\n\n");
    qnd_destroy(ctx);
    free(Psi);
    free(Eobs);
    free(Aobs);
    free(codes);
    free(dims);
    Finalize_parameters(&P);
    Finalize_parameters(&P_imag);
    Finalize_field_vectors(&F);
    return 0;
}
