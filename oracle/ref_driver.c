/*
 * Test-infrastructure driver for the reference build under oracle/_ref.
 *
 * The reference's own main (src/kssd3.c) and top-level dispatcher
 * (src/global_wrapper.c:229-272) include a header generated by the
 * reference's release tooling, so they are left out of this build.  This
 * driver replaces only that dispatch: it hands argv to the reference's own
 * subcommand entry points (cmd_ani, src/command_ani_wrapper.c:1504;
 * cmd_matrix, src/command_matrix_wrapper.c:552; cmd_sketch,
 * src/command_sketch_wrapper.c:1045) exactly as argp would after consuming
 * the subcommand word.  It is a checker, never part of the product.
 */
#include <argp.h>
#include <stdio.h>
#include <string.h>

int cmd_ani(struct argp_state *state);
int cmd_matrix(struct argp_state *state);
int cmd_sketch(struct argp_state *state);

const char *argp_program_version = "kssd3a-reference-oracle";

int main(int argc, char **argv)
{
	if (argc < 2) {
		fprintf(stderr, "usage: %s ani|matrix|sketch [args...]\n", argv[0]);
		return 64;
	}
	struct argp_state st;
	memset(&st, 0, sizeof st);
	st.argc = argc;
	st.argv = argv;
	st.next = 2; /* argv[1] (the subcommand) has been consumed */
	st.name = argv[0];
	st.input = NULL;
	if (strcmp(argv[1], "ani") == 0)
		{
		cmd_ani(&st);
		return 0;
	}
	if (strcmp(argv[1], "matrix") == 0) {
		cmd_matrix(&st);
		return 0;
	}
	if (strcmp(argv[1], "sketch") == 0) {
		cmd_sketch(&st);
		return 0;
	}
	fprintf(stderr, "%s: unsupported subcommand %s\n", argv[0], argv[1]);
	return 64;
}
