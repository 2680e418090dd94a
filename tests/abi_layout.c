/* Prints sizeof/offsetof of the C-ABI structs (compiled by gcc in
 * test_abi.py to check the ctypes mirrors in warpcfd_amd/_lib.py). */
#include <stddef.h>
#include <stdio.h>
#include "wcfd.h"
#define P(T, F) printf(#T "." #F " %zu\n", offsetof(T, F))
int main(void) {
  printf("wcfd_mesh %zu\n", sizeof(wcfd_mesh));
  P(wcfd_mesh, n_slots); P(wcfd_mesh, owner); P(wcfd_mesh, col); P(wcfd_mesh, slot_sn);
  printf("wcfd_bc %zu\n", sizeof(wcfd_bc));
  printf("wcfd_solve_info %zu\n", sizeof(wcfd_solve_info));
  P(wcfd_solve_info, residual);
  printf("wcfd_simple_params %zu\n", sizeof(wcfd_simple_params));
  P(wcfd_simple_params, atol); P(wcfd_simple_params, rtol0); P(wcfd_simple_params, maxiter); P(wcfd_simple_params, first_iteration); P(wcfd_simple_params, check_every_u); P(wcfd_simple_params, p_refine); P(wcfd_simple_params, graph_groups_p); P(wcfd_simple_params, maxiter_u);
  printf("wcfd_simple_state %zu\n", sizeof(wcfd_simple_state));
  P(wcfd_simple_state, resid);
  printf("wcfd_simple_info %zu\n", sizeof(wcfd_simple_info));
  P(wcfd_simple_info, continuity); P(wcfd_simple_info, velocity_change);
  printf("wcfd_batch %zu\n", sizeof(wcfd_batch));
  P(wcfd_batch, cell_off); P(wcfd_batch, active); P(wcfd_batch, stop); P(wcfd_batch, monitor); P(wcfd_batch, outer_iters); P(wcfd_batch, max_case_blocks); P(wcfd_batch, n_bface_blk); P(wcfd_batch, cell_blk); P(wcfd_batch, bface_blk);
  return 0;
}
