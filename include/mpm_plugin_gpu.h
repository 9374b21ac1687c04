/*
 * mpm_plugin_gpu.h -- optional extra symbol a material or boundary-condition
 * plugin may export so that the GPU step evaluates it on the device.
 *
 * The reference loads plugins with dlopen/dlsym (src/main.c:427-456 for
 * materials, 500-543 for boundary conditions) and calls them back with job_t*
 * on the host.  That interface is unchanged.  A plugin that also exports
 *     int material_gpu_descriptor(const job_t *, mpm_material_gpu_desc *)
 * or
 *     int bc_gpu_descriptor(const job_t *, mpm_bc_gpu_desc *)
 * tells the drop-in step which device implementation reproduces it and with
 * which parameters.  The drop-in looks the symbol up in the same shared
 * object as the plugin's calculate_stress_threaded / bc_time_varying
 * (dladdr + dlsym).  Plugins without it still work: materials then run on the
 * host between the G2P and P2G kernels, and boundary conditions run their
 * bc_time_varying on the host with the resulting masks uploaded when they
 * change.
 *
 * Both functions are called after material_init / bc_validate, return 1 when
 * they filled the descriptor and 0 otherwise.
 */
#ifndef MPM_PLUGIN_GPU_H
#define MPM_PLUGIN_GPU_H

#include "mpm_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_PLUGIN_GPU_ABI 1

enum mpm_gpu_material_kind {
    MPM_GPU_MATERIAL_G_LOCAL_MU2 = 1,  /* materialsrc/g_local_mu2.c */
    MPM_GPU_MATERIAL_ISOLIN = 2,       /* materialsrc/isolin.c */
    MPM_GPU_MATERIAL_DP_RI_MT = 3      /* materialsrc/dp_ri_mt.c */
};

typedef struct mpm_material_gpu_desc {
    int abi;        /* MPM_PLUGIN_GPU_ABI */
    int kind;       /* mpm_gpu_material_kind */
    /* g_local_mu2: E, nu, mu_s, mu_2, I_0, grains_d, grains_rho, rho_crit;
     * isolin: E, nu; dp_ri_mt: E, nu, mu_s */
    double params[8];
} mpm_material_gpu_desc;

enum mpm_gpu_bc_kind {
    MPM_GPU_BC_BOX = 1,          /* bcsrc/box.c */
    MPM_GPU_BC_SQUARE_SILO = 2   /* bcsrc/square_silo.c */
};

typedef struct mpm_bc_gpu_desc {
    int abi;        /* MPM_PLUGIN_GPU_ABI */
    int kind;       /* mpm_gpu_bc_kind */
    double fp64[4]; /* the plugin's fp64 properties */
    int i32[4];     /* the plugin's integer properties */
} mpm_bc_gpu_desc;

typedef int (*mpm_material_gpu_descriptor_fn)(const job_t *, mpm_material_gpu_desc *);
typedef int (*mpm_bc_gpu_descriptor_fn)(const job_t *, mpm_bc_gpu_desc *);

#ifdef __cplusplus
}
#endif
#endif /* MPM_PLUGIN_GPU_H */
