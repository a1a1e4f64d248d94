/* Drives the CPU oracle (oracle/sbe_oracle.c) through every entry point on a
 * small two-band chain, for the AddressSanitizer / UndefinedBehaviorSanitizer
 * build of tests/test_sanitizers.py.  Exit status 0 when every call succeeds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/sbe_oracle.h"

int main(void)
{
    enum { NC = 5, N = 2 };
    int cells[NC * 3] = {-2, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 0, 2, 0, 0};
    double H[NC * N * N * 2] = {0}, D[NC * 3 * N * N * 2] = {0};
    for (int c = 0; c < NC; c++) {
        int R = cells[3 * c];
        double a = 0.02 * exp(-fabs((double)R)), ph = 0.3 * R;
        /* H(R)[0][1] = a e^{i ph}, H(-R) = H(R)^dagger, onsite split */
        H[((c * N + 0) * N + 1) * 2] = a * cos(ph);
        H[((c * N + 0) * N + 1) * 2 + 1] = a * sin(ph);
        H[((c * N + 1) * N + 0) * 2] = a * cos(ph);
        H[((c * N + 1) * N + 0) * 2 + 1] = -a * sin(ph);
        if (R == 0) {
            H[((c * N + 0) * N + 1) * 2 + 1] = 0;
            H[((c * N + 1) * N + 0) * 2 + 1] = 0;
            H[((c * N + 1) * N + 1) * 2] = 0.1;
            D[(((c * 3 + 0) * N + 0) * N + 1) * 2] = 2.0;
            D[(((c * 3 + 0) * N + 1) * N + 0) * 2] = 2.0;
        }
    }
    sbe_model m = {N, NC, cells, H, D, {5.0, 0, 0, 0, 5.0, 0, 0, 0, 5.0}};
    sbe_grid g = {1, {16, 1, 1}, {8, 1, 1}, {0, 0, 0}};
    sbe_params p = {0.002, 0.001, 0.001, 1e-12, 1e-8, 6, 0};
    sbe_subpulse pu = {SBE_PULSE_SIN2, 3, 0.002, 0.03, 0.0, 150.0, -1, 1, 0, 1, {1, 0, 0}};
    oracle_ctx *c = oracle_create(&m, &g, &p, &pu, 1);
    if (!c) return 1;
    double mu;
    if (oracle_fermi_level(c, 1, 0, &mu)) return 2;
    long S = oracle_state_size(c);
    double *x = calloc(2 * S, sizeof(double)), *dx = calloc(2 * S, sizeof(double));
    if (oracle_initial_state(c, mu, x)) return 3;
    if (oracle_rhs(c, x, dx, -10.0)) return 4;
    double j[3];
    if (oracle_observe_j(c, x, j)) return 5;
    sbe_region reg[2] = {{{0, 0, 0}, INFINITY, 1}, {{0.25, 0, 0}, 0.2, 0}};
    double ws[2], jr[6], oh[4], ow[4], mn[2], mx[2], wan[2 * 4 * 2];
    if (oracle_observe_full(c, x, reg, 2, 0.001, ws, jr, oh, ow, mn, mx, wan)) return 6;
    double t[9], jo[27];
    long rhs = 0;
    long steps = oracle_integrate(c, x, 9, t, jo, NULL, NULL, &rhs);
    if (steps <= 0) return 7;
    printf("oracle ok: %ld steps, %ld rhs, j %.3e\n", steps, rhs, jo[3 * 8]);
    free(x);
    free(dx);
    oracle_destroy(c);
    return 0;
}
