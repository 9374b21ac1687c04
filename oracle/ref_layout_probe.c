/* Layout probe for the reference's plugin-ABI data types.
 *
 * Test infrastructure: compiled only by oracle/Makefile against the
 * reference's own headers (libmpm/particle.h, node.h, element.h), which
 * include nothing that this image lacks.  It prints sizeof/offsetof for
 * every field that the drop-in header include/mpm_abi.h must reproduce.
 */
#include <stdio.h>
#include <stddef.h>
#include "particle.h"
#include "node.h"
#include "element.h"

#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f))

int main(void)
{
    printf("particle_t sizeof %zu\n", sizeof(particle_t));
    P(particle_t, x); P(particle_t, y); P(particle_t, xl); P(particle_t, yl);
    P(particle_t, v); P(particle_t, v0); P(particle_t, s); P(particle_t, m);
    P(particle_t, x_t); P(particle_t, y_t); P(particle_t, x_tt); P(particle_t, y_tt);
    P(particle_t, bx); P(particle_t, by); P(particle_t, sxx); P(particle_t, sxy);
    P(particle_t, syy); P(particle_t, T); P(particle_t, exx_t); P(particle_t, exy_t);
    P(particle_t, eyy_t); P(particle_t, wxy_t); P(particle_t, L); P(particle_t, Fxx);
    P(particle_t, Fxy); P(particle_t, Fyx); P(particle_t, Fyy); P(particle_t, ux);
    P(particle_t, uy); P(particle_t, color); P(particle_t, state); P(particle_t, material);
    P(particle_t, grad_s); P(particle_t, element); P(particle_t, real_sxx);
    P(particle_t, real_sxy); P(particle_t, real_syy); P(particle_t, real_state);
    P(particle_t, r1_initial); P(particle_t, r2_initial); P(particle_t, r1);
    P(particle_t, r2); P(particle_t, corners); P(particle_t, cornersl); P(particle_t, sc);
    P(particle_t, grad_sc); P(particle_t, corner_elements); P(particle_t, material_data);
    P(particle_t, id);
    printf("node_t sizeof %zu\n", sizeof(node_t));
    P(node_t, num_filled_element_neighbors); P(node_t, mass_filled_element_neighbors);
    P(node_t, m); P(node_t, x); P(node_t, y); P(node_t, ux); P(node_t, uy);
    P(node_t, x_t); P(node_t, y_t); P(node_t, x_tt); P(node_t, y_tt);
    P(node_t, mx_t); P(node_t, my_t); P(node_t, mx_tt); P(node_t, my_tt);
    P(node_t, fx); P(node_t, fy); P(node_t, rho); P(node_t, velocity_update_flag);
    printf("element_t sizeof %zu\n", sizeof(element_t));
    P(element_t, color); P(element_t, color_idx); P(element_t, nodes); P(element_t, n);
    P(element_t, m); P(element_t, filled); P(element_t, neighbors); P(element_t, grad_x);
    P(element_t, grad_y); P(element_t, grad_mag); P(element_t, n_x); P(element_t, n_y);
    P(element_t, n_theta); P(element_t, v); P(element_t, p_index); P(element_t, p_dist);
    P(element_t, sxx); P(element_t, sxy); P(element_t, syy); P(element_t, exx_t);
    P(element_t, exy_t); P(element_t, eyy_t); P(element_t, wxy_t); P(element_t, p);
    P(element_t, tau); P(element_t, kku_element); P(element_t, f_element);
    P(element_t, jacobian);
    return 0;
}
