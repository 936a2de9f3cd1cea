/* Prints the sizes and field offsets of the C ABI structs, so the ctypes
 * mirrors in collision_amd/_lib.py can be checked against the header. */
#include <stddef.h>
#include <stdio.h>

#include "collision_amd.h"

#define F(T, m) printf("%s.%s %zu\n", #T, #m, offsetof(T, m))
int main(void) {
    printf("cs_params %zu\ncs_mesh_view %zu\ncs_pass_info %zu\ncs_step_report %zu\n",
           sizeof(cs_params), sizeof(cs_mesh_view), sizeof(cs_pass_info), sizeof(cs_step_report));
    F(cs_params, max_ccd_passes); F(cs_params, device); F(cs_params, force_full_passes);
    F(cs_params, max_strain_sweeps); F(cs_params, log_pairs); F(cs_params, pass_timing);
    F(cs_mesh_view, tris); F(cs_mesh_view, is_fixed);
    F(cs_pass_info, candidates); F(cs_pass_info, tested_pairs); F(cs_pass_info, ms_narrow_kernel);
    F(cs_step_report, candidates); F(cs_step_report, ms_strain);
    F(cs_step_report, strain_levels); F(cs_step_report, strain_slots);
    return 0;
}
