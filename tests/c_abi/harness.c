/*
 * C stand-in for the Rust caller of the device solve (src/solver.rs:1778-1808):
 * packs a small system in the reference layout, calls vx_solve_pcg through
 * the C ABI, and on a nonzero status takes the fallback branch (a reference
 * CPU PCG written here in plain C) exactly as solve_coupled does on Err.
 * Prints one line: "device <iterations> <residual>" or "fallback <status> <message>",
 * then the solution, and exits 0 when the solution matches the known answer.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "vx_solver.h"

/* Two coupled REVs: diagonal blocks 2 I, one off-diagonal block -0.2 I
 * (exact solution x_i = (2 b_i + 0.2 b_j) / 3.96). */
#define N 2
static const float rhs[N * 3] = {1.0f, 2.0f, 3.0f, -1.0f, 0.5f, 4.0f};

static void pack(uint32_t nbr[N * 6], float diag[N * 9], float blocks[N * 54], float inv[N * 9]) {
    memset(blocks, 0, sizeof(float) * N * 54);
    memset(diag, 0, sizeof(float) * N * 9);
    memset(inv, 0, sizeof(float) * N * 9);
    for (int i = 0; i < N; ++i) {
        for (int s = 0; s < 6; ++s) nbr[i * 6 + s] = VX_EMPTY;
        nbr[i * 6 + 0] = (uint32_t)(1 - i);
        for (int d = 0; d < 3; ++d) {
            diag[i * 9 + d * 4] = 2.0f;
            inv[i * 9 + d * 4] = 0.5f;
            blocks[i * 54 + d * 4] = -0.2f;
        }
    }
}

/* Fallback: plain block-Jacobi PCG on the same rows (the caller's CPU path). */
static int cpu_pcg(const uint32_t *nbr, const float *diag, const float *blocks, const float *inv, float *x) {
    double r[N * 3], z[N * 3], p[N * 3], ap[N * 3], rho = 0.0, bb = 0.0;
    for (int k = 0; k < N * 3; ++k) {
        x[k] = 0.0f;
        r[k] = rhs[k];
        bb += r[k] * r[k];
    }
    for (int it = 0; it < 100; ++it) {
        double rn = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < 3; ++a) {
                z[i * 3 + a] = 0.0;
                for (int b = 0; b < 3; ++b) z[i * 3 + a] += inv[i * 9 + a * 3 + b] * r[i * 3 + b];
            }
        for (int k = 0; k < N * 3; ++k) rn += r[k] * z[k];
        for (int k = 0; k < N * 3; ++k) p[k] = z[k] + (it ? rn / rho : 0.0) * p[k];
        rho = rn;
        double pap = 0.0;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < 3; ++a) {
                double y = 0.0;
                for (int b = 0; b < 3; ++b) y += diag[i * 9 + a * 3 + b] * p[i * 3 + b];
                for (int s = 0; s < 6; ++s) {
                    uint32_t j = nbr[i * 6 + s];
                    if (j == VX_EMPTY) continue;
                    for (int b = 0; b < 3; ++b) y += blocks[i * 54 + s * 9 + a * 3 + b] * p[j * 3 + b];
                }
                ap[i * 3 + a] = y;
                pap += p[i * 3 + a] * y;
            }
        const double alpha = rho / pap;
        double rr = 0.0;
        for (int k = 0; k < N * 3; ++k) {
            x[k] += (float)(alpha * p[k]);
            r[k] -= alpha * ap[k];
            rr += r[k] * r[k];
        }
        if (sqrt(rr / bb) < 1e-6) return it + 1;
    }
    return -1;
}

int main(void) {
    uint32_t nbr[N * 6];
    float diag[N * 9], blocks[N * 54], inv[N * 9], dx[N * 3];
    pack(nbr, diag, blocks, inv);
    uint32_t iterations = 0;
    float residual = 0.0f;
    int32_t status = vx_solve_pcg(rhs, nbr, diag, blocks, inv, N, 80, 1e-5f, dx, &iterations, &residual);
    if (status == VX_OK) {
        printf("device %u %.3e\n", iterations, residual);
    } else {
        printf("fallback %d %s\n", status, vx_last_error());
        if (cpu_pcg(nbr, diag, blocks, inv, dx) < 0) return 2;
    }
    int ok = 1;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            const double expect = (2.0 * rhs[i * 3 + a] + 0.2 * rhs[(1 - i) * 3 + a]) / 3.96;
            printf("%.6f ", dx[i * 3 + a]);
            if (fabs(dx[i * 3 + a] - expect) > 1e-5) ok = 0;
        }
    printf("\n");

    /* collision::project from C: a voxel jumps three cells through a fixed
       one (src/collision.rs:731-747); the fallback is the caller's CPU path. */
    float pos[6] = {-0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f};
    float prev[6], rest[6], mass[2] = {6e-4f, 6e-4f};
    uint8_t kinematic[2] = {0, 1};
    uint32_t bonds[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    for (int k = 0; k < 6; ++k) prev[k] = rest[k] = pos[k];
    const float wall = pos[3];
    pos[0] += 3.0f * 0.01f;
    vx_contact contacts[4];
    vx_contact_report report;
    status = vx_contact_project(2, pos, prev, rest, mass, kinematic, NULL, bonds, contacts, 4, &report);
    if (status == VX_OK) {
        const int held = report.contacts >= 1 && pos[3] == wall && wall - pos[0] >= 0.01f - 1e-6f &&
                         contacts[0].i == 0 && contacts[0].j == 1;
        printf("contact %s %llu\n", held ? "ok" : "wrong", (unsigned long long)report.contacts);
        if (!held) ok = 0;
    } else {
        printf("contact fallback %d\n", status);
    }
    return ok ? 0 : 1;
}
