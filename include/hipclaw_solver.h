/*
 * hipclaw_solver: Riemann solvers supplied by an application.
 *
 * The reference lets an application plug its own device Riemann solvers into
 * fc2d_cudaclaw through the vtable slots cuda_rpn2 / cuda_rpt2
 * (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw.h:159-162), device function
 * pointers of the types cudaclaw_cuda_rpn2_t / cudaclaw_cuda_rpt2_t
 * (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw_cuda.h:57-67) that every interface
 * of the batched update calls indirectly.
 *
 * Here a solver is a compile-time functor (rpn2 and rpt2 as inlined member
 * templates) that the application instantiates into the step kernel itself,
 * by compiling one HIP source with forestclaw_amd/csrc/hipclaw_user_solver.h
 * (the functor contract is documented there).  That yields the descriptor
 * below; hipclaw_register_solver turns it into a solver number for
 * hipclaw_config_t.solver, and fc2d_hipclaw_vtable_t.rp_solver does the same
 * for the fc2d_hipclaw plugin.
 */
#ifndef HIPCLAW_SOLVER_H
#define HIPCLAW_SOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Version of the descriptor and of the kernel argument block its launchers
 * take; a descriptor built against another version is refused. */
#define HIPCLAW_SOLVER_ABI 1

/* Registered solvers get numbers from HIPCLAW_RP_USER on. */
#define HIPCLAW_RP_USER 16
#define HIPCLAW_MAX_USER_SOLVERS 16

/* Launches one batched step of the solver's step kernel.  tile: 8 or 16;
 * waves: waves per workgroup of a 16 x 16 tile (0 = default); step_args: the
 * library's kernel argument block (step_args_size bytes); stream, ev_start,
 * ev_stop: a hipStream_t and two hipEvent_t (may be null).  Returns a
 * hipError_t value. */
typedef int (*hipclaw_step_launcher_t)(int tile, int waves, const void *step_args, void *stream,
                                       void *ev_start, void *ev_stop);

typedef struct hipclaw_solver {
    int abi;                 /* HIPCLAW_SOLVER_ABI */
    int step_args_size;      /* size of the argument block the launchers were built for */
    const char *name;
    int meqn, mwaves, maux;
    int conservative;        /* amdq + apdq = f(qR) - f(qL): the FP32-bulk flux form */
    int time_sync;           /* provides the conservative flux the time-sync registers keep */
    int tiles;               /* tile sizes instantiated: bit 3 = 8 x 8, bit 4 = 16 x 16 */
    hipclaw_step_launcher_t launch_fp64;   /* the FP64 parity form */
    hipclaw_step_launcher_t launch_fp32;   /* the FP32-bulk form, or null */
} hipclaw_solver_t;

/* Register a solver; *id receives its number for hipclaw_config_t.solver.
 * Registering the same descriptor again returns the same number.  The
 * descriptor must stay valid while the library is in use. */
int hipclaw_register_solver(const hipclaw_solver_t *s, int *id);

/* The descriptor of solver `id` (built-in solvers 0..2 have none: null). */
const hipclaw_solver_t *hipclaw_registered_solver(int id);

#ifdef __cplusplus
}
#endif
#endif
