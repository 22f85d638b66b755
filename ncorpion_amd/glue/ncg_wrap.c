/*
 * ncg_wrap.c -- link-time redirection of NcorpiON's hot-path calls.
 *
 * Linked with -Wl,--wrap=vector_field,--wrap=brute_force,--wrap=mesh,
 * --wrap=get_neighbours_mesh,--wrap=display: every call from rk4.o
 * (rk4.c:85, 593, 700, 741, 762, 767, 776, 806, 876, 909; display:
 * rk4.c:506, 570, 643, 692, 745, 812, 869, 902, 938) lands here and goes to
 * the GPU glue.
 */
struct moonlet;

void ncg_dropin_vector_field(struct moonlet *X);
void ncg_dropin_brute_force(struct moonlet *X);
void ncg_dropin_mesh(struct moonlet *X);
void ncg_dropin_get_neighbours_mesh(struct moonlet *X);
void ncg_dropin_display(struct moonlet *X);

void __wrap_vector_field(struct moonlet *X) { ncg_dropin_vector_field(X); }
void __wrap_brute_force(struct moonlet *X) { ncg_dropin_brute_force(X); }
void __wrap_mesh(struct moonlet *X) { ncg_dropin_mesh(X); }
void __wrap_get_neighbours_mesh(struct moonlet *X) { ncg_dropin_get_neighbours_mesh(X); }
void __wrap_display(struct moonlet *X) { ncg_dropin_display(X); }
