#!/usr/bin/env python3
"""Writes the reference driver with the HIP back-end dropped in -- TEST
INFRASTRUCTURE ONLY (run by build_dropin.sh).

Reads swe_code/main/main.c from the reference tree, inserts the HIP-side
lines of the integration (the third back-end branch, next to the OpenCL and
CPU ones of main.c:338-356) at six anchor lines, and writes the result to
the given path (a temporary directory; no reference text is kept in this
repository).  The inserted code is this file's own; the anchors are the
reference lines it is inserted after or before.

Run time: SWE_USE_HIP=0 (default) keeps the reference's CPU/OpenCL path;
1 replaces RK_substep by swe_hip_rk_substep (four calls per step, as the
reference); 2 runs the whole step loop as one fused swe_hip_rk_steps call;
SWE_HIP_EXACT=1 selects the bitwise-reference arithmetic.  A header built
with RUN_32BIT (the reference's shipped FP32 type) uses the *32 entry
points.
"""
import sys

INCLUDE = r'''
#include <stdio.h>
#include <stdlib.h>
#include <swe_hip.h>
static int swe_use_hip(void)
{
    const char *v = getenv("SWE_USE_HIP");
    return v ? atoi(v) : 0;
}
static void hip_check(int rc, const char *what)
{
    if (rc != SWE_HIP_OK) {
        printf("Error in %s: %s: %s\n", what, swe_hip_strerror(rc), swe_hip_last_error());
        exit(-1);
    }
}
'''

SETUP = r'''
    /* HIP back-end set-up (replaces the OpenCL set-up above) */
    const int use_hip = swe_use_hip();
    swe_hip_ctx *hip = NULL;
    if (use_hip) {
        const char *ex = getenv("SWE_HIP_EXACT");
        const unsigned hip_flags = (ex && atoi(ex)) ? SWE_HIP_EXACT : 0u;
#ifdef RUN_32BIT
        swe_hip_static32 hs;
#else
        swe_hip_static hs;
#endif
        hs.compute_size = LPSMD->compute_size;
        hs.patch_size = LPSMD->patch_size;
        hs.compute_pid_s = LPSMD->compute_pid_s;
        hs.Nnbr = LPSMD->Nnbr;
        hs.padded_Nnbr = LPSMD->padded_Nnbr;
#ifdef USE_CFDL
        hs.cfdl = 1;
#else
        hs.cfdl = 0;
#endif
#ifdef USE_SFDL
        hs.sfdl = 1;
#else
        hs.sfdl = 0;
#endif
        hs.simd = SIMD_LENGTH;
        hs.gh0 = LPSMD->gh0;
        hs.dt = LPSMD->dt;
        hs.x = LPSMD->x; hs.y = LPSMD->y; hs.z = LPSMD->z; hs.f = LPSMD->f; hs.ghm = LPSMD->ghm;
        hs.p_u = LPSMD->p_u; hs.p_v = LPSMD->p_v; hs.p_w = LPSMD->p_w;
        hs.gradghm = LPSMD->gradghm;
        hs.idx = LPSMD->idx;
        hs.Dx = LPSMD->Dx; hs.Dy = LPSMD->Dy; hs.Dz = LPSMD->Dz; hs.L = LPSMD->L;
        const double hip_t0 = getTime();
#ifdef RUN_32BIT
        hip_check(swe_hip_create32(&hs, 1, hip_flags, 0, &hip), "swe_hip_create32");
#else
        hip_check(swe_hip_create(&hs, 1, hip_flags, 0, &hip), "swe_hip_create");
#endif
        local_timer.t_ocl = getTime() - hip_t0;
        if (mpi_rank == 0)
            printf("HIP back-end: %s%s\n", use_hip == 2 ? "fused steps" : "RK_substep drop-in",
                   hip_flags ? " (exact)" : "");
    }
'''

SET_STATE = r'''
        if (use_hip) {
#ifdef RUN_32BIT
            hip_check(swe_hip_set_state32(hip, 0, H), "swe_hip_set_state32");
#else
            hip_check(swe_hip_set_state(hip, 0, H), "swe_hip_set_state");
#endif
            hip_check(swe_hip_set_timing(hip, use_hip == 1), "swe_hip_set_timing");
        }
'''

SUBSTEP = r'''                else if (use_hip == 1) hip_check(swe_hip_rk_substep(hip, rk_val), "swe_hip_rk_substep");
                else if (use_hip == 2) {
                    if (i == 1 && rk_val == 0) {
                        hip_check(swe_hip_rk_steps(hip, sim_params.nsteps), "swe_hip_rk_steps");
                        hip_check(swe_hip_synchronize(hip), "swe_hip_synchronize");
                    }
                }
'''

GET_STATE = r'''
        if (use_hip) {
            /* read-back (the OpenCL path maps H_buff here) and the phase timers */
#ifdef RUN_32BIT
            hip_check(swe_hip_get_state32(hip, 0, H), "swe_hip_get_state32");
#else
            hip_check(swe_hip_get_state(hip, 0, H), "swe_hip_get_state");
#endif
            swe_hip_timers ht;
            hip_check(swe_hip_get_timers(hip, &ht, 1), "swe_hip_get_timers");
            local_timer.t_eval_rhs[attempt] += ht.t_eval_rhs;
            local_timer.t_eval_K[attempt] += ht.t_eval_K;
            local_timer.t_update_D[attempt] += ht.t_update_D;
            local_timer.t_update_H[attempt] += ht.t_update_H;
        }
'''

RELEASE = r'''
    if (use_hip)
        swe_hip_destroy(hip);
'''

# (anchor, where, text): inserted once, after or before the first line whose
# stripped text starts with the anchor.
EDITS = [
    ("#include <profiling.h>", "after", INCLUDE),
    ("for (int attempt = 0; attempt < sim_params.nattempts; attempt++)", "before", SETUP),
    ("copy_fp_arr(H, H_init, LPSMD->patch_size * 4);", "after", SET_STATE),
    ("else RK_substep(H, K, F, D, rk_val);", "replace_else", SUBSTEP),
    ("// Compare H array to expected values at t=100", "before", GET_STATE),
    ("process_profiling_results();", "before", RELEASE),
]


def patch(text: str) -> str:
    lines = text.split("\n")
    for anchor, where, ins in EDITS:
        hits = [k for k, ln in enumerate(lines) if ln.strip().startswith(anchor)]
        if len(hits) < 1:
            raise SystemExit(f"dropin_main: anchor not found: {anchor!r}")
        k = hits[0]
        block = ins.strip("\n").split("\n")
        if where == "after":
            lines[k + 1:k + 1] = block
        elif where == "before":
            lines[k:k] = block
        else:   # the reference's CPU branch stays last: else if (hip) ... else RK_substep
            indent = lines[k][: len(lines[k]) - len(lines[k].lstrip())]
            lines[k:k + 1] = block + [indent + lines[k].strip()]
    return "\n".join(lines)


def main():
    src, dst = sys.argv[1], sys.argv[2]
    with open(src) as fp:
        text = fp.read()
    with open(dst, "w") as fp:
        fp.write(patch(text))


if __name__ == "__main__":
    main()
