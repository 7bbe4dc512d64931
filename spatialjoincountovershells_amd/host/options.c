/*
 * options.c -- argv handling of the drop-in CLI.
 *
 * Same semantics as the reference's option scanner (src/GetOption/
 * GetOption.c:7-110) and option block of astr_rcount's main()
 * (src/astr_rcount/main.c:492-626): exact-length matches, the last
 * occurrence wins, a file list ends at the next argument starting with '-',
 * `--radius a:b` takes K from the text after the last ':', and only the
 * first validation error is reported.  The messages are the reference's,
 * byte for byte, because scripts match them.
 */
#include <stdlib.h>
#include <string.h>

#include "sjcs_host.h"

static int opt_find(int argc, char **argv, const char *opt, int offset,
                    int (*store)(const char *, void *), void *dst)
{
    size_t len = strlen(opt);
    int i;
    for (i = 1; i < argc; i++) {
        if (strlen(argv[i]) != len || memcmp(argv[i], opt, len) != 0)
            continue;
        i += offset + 1;
        if (i >= argc)
            return -1;
        store(argv[i], dst);
    }
    return 0;
}

static int store_string(const char *v, void *dst)
{
    *(char **)dst = (char *)v;
    return 0;
}

static int store_int(const char *v, void *dst)
{
    *(int *)dst = (int)strtol(v, NULL, 10);
    return 0;
}

int sjcs_opt_string(int argc, char **argv, const char *opt, int offset, char **dst)
{
    return opt_find(argc, argv, opt, offset, store_string, dst);
}

int sjcs_opt_int(int argc, char **argv, const char *opt, int offset, int *dst)
{
    return opt_find(argc, argv, opt, offset, store_int, dst);
}

void sjcs_opt_range(int argc, char **argv, const char *opt, int *beg, int *end)
{
    size_t len = strlen(opt);
    int i, j;
    for (i = 1; i < argc; i++) {
        if (strlen(argv[i]) != len || memcmp(argv[i], opt, len) != 0)
            continue;
        for (j = i + 1; j < argc && argv[j][0] != '-'; j++)
            ;
        if (j != i + 1) {
            *beg = i + 1;
            *end = j;
        }
    }
}

int sjcs_opt_flag(int argc, char **argv, const char *opt)
{
    size_t len = strlen(opt);
    int i;
    for (i = 1; i < argc; i++)
        if (strlen(argv[i]) == len && memcmp(argv[i], opt, len) == 0)
            return 1;
    return 0;
}

static void first_error(sjcs_cli *cli, const char *msg)
{
    if (!cli->errflag) {
        cli->errflag = 1;
        strncpy(cli->errmsg, msg, SJCS_ERRMSG_LEN - 1);
        cli->errmsg[SJCS_ERRMSG_LEN - 1] = '\0';
    }
}

void sjcs_cli_parse(int argc, char **argv, sjcs_cli *cli)
{
    memset(cli, 0, sizeof(*cli));
    cli->particle_beg = -1;
    cli->particle_end = -1;
    cli->nthreads = -1;
    cli->nsetups = -1;
    cli->first_nbacklogs = 1;
    cli->first_nsetups = 1;

    sjcs_opt_string(argc, argv, "--output-file", 0, &cli->outfile);
    sjcs_opt_range(argc, argv, "--particle-files", &cli->particle_beg, &cli->particle_end);
    cli->compressed = sjcs_opt_flag(argc, argv, "--compressed");
    sjcs_opt_string(argc, argv, "--halo-file", 0, &cli->halo_file);
    sjcs_opt_string(argc, argv, "--filter-file", 0, &cli->filter_file);

    sjcs_opt_string(argc, argv, "--radius", 0, &cli->range_str);
    if (cli->range_str != NULL) {
        char *d;
        cli->minrad = strtod(cli->range_str, NULL);
        if ((d = strchr(cli->range_str, ':')) != NULL) {
            cli->maxrad = strtod(d + 1, NULL);
            if ((d = strrchr(cli->range_str, ':')) != NULL)
                cli->nrads = (int)strtol(d + 1, NULL, 10);
        }
    }
    if (cli->range_str == NULL || cli->nrads <= 0 || cli->minrad <= 0 || cli->maxrad <= 0)
        first_error(cli, "# of ranges, minimum range, maximum range must be > 0.\n"
                         "      use --radius and fix.\n");
    else if (cli->minrad == cli->maxrad)
        first_error(cli, "minimum range must be <  maximum range.\n"
                         "      use --radius and fix.\n");
    if (cli->minrad > cli->maxrad) {
        double t = cli->minrad;
        cli->minrad = cli->maxrad;
        cli->maxrad = t;
    }

    sjcs_opt_int(argc, argv, "--backlog", 0, &cli->maxbacklogs);
    if (cli->maxbacklogs < 0)
        first_error(cli, "max # of backlog indexes must be >= 0.\n"
                         "      use --backlog and fix.\n");

    sjcs_opt_int(argc, argv, "--thread", 0, &cli->nthreads);
    sjcs_opt_int(argc, argv, "--next-setup", 0, &cli->nsetups);
    if (cli->nthreads < 1)
        first_error(cli, "# of threads must be >= 1.\n"
                         "      use --thread and fix.\n");
    if (cli->nsetups < 1)
        first_error(cli, "# of backlog index construction threads must be >= 1.\n"
                         "      use --next-setup and fix.\n");

    sjcs_opt_string(argc, argv, "--first-setup", 0, &cli->setup_str);
    if (cli->setup_str != NULL) {
        char *d;
        cli->first_nbacklogs = (int)strtol(cli->setup_str, NULL, 10);
        if ((d = strchr(cli->setup_str, ':')) != NULL)
            cli->first_nsetups = (int)strtol(d + 1, NULL, 10);
    }

    if (cli->maxbacklogs == 0) {
        if (cli->nthreads < cli->nsetups)
            first_error(cli, "in noBacklogExecution; # of threads must be"
                             ">= # of index construction threads.\n"
                             "      use --thread, --next-setup and fix.\n");
    } else {
        if (cli->nthreads < cli->nsetups + 1 || cli->nthreads < cli->first_nsetups + 1)
            first_error(cli, "in withBacklogExecution; # of threads must be"
                             "> # of index construction threads.\n"
                             "      use --thread, --next-setup, --first-setup and fix.\n");
        if (cli->maxbacklogs < cli->first_nbacklogs)
            first_error(cli, "in withBacklogExecution; max # of backlog indexes must be"
                             "> # of first backlog indexes.\n"
                             "      use --backlog, --first-setup and fix.\n");
    }

    if (cli->particle_beg <= 0)
        first_error(cli, "particle file is missing.\n"
                         "      use --particle-files and fix.\n");
    if (cli->halo_file == NULL)
        first_error(cli, "halo file is missing.\n"
                         "      use --halo-file and fix.\n");
}

static const char *or_null(const char *s)
{
    return s ? s : "(null)";
}

void sjcs_cli_print_info(FILE *fp, const sjcs_cli *cli, char **argv)
{
    int i;
    fputs("\n--execution info--\n", fp);
    fprintf(fp, "  particle files (compressed = %s): \n", cli->compressed ? "true" : "false");
    for (i = cli->particle_beg; i < cli->particle_end; i++)
        fprintf(fp, "   |- %s\n", argv[i]);
    fprintf(fp,
            "  halo file = %s\n"
            "  halo filter file = %s\n"
            "  # of threads in thread pool = %d\n"
            "  # of index construction threads = %d\n"
            "  # of radiuses = %d\n"
            "   |- minimum = %f, maximum = %f\n"
            "  max # of backlog indexes = %d\n"
            "  %s# of first backlog indexes = %d\n"
            "  %s# of first index construction threads = %d\n",
            or_null(cli->halo_file), or_null(cli->filter_file), cli->nthreads, cli->nsetups,
            cli->nrads, cli->minrad, cli->maxrad, cli->maxbacklogs,
            cli->maxbacklogs == 0 ? "! ignored ! " : "", cli->first_nbacklogs,
            cli->maxbacklogs == 0 ? "! ignored ! " : "", cli->first_nsetups);
}

void sjcs_cli_print_usage(FILE *fp, const char *argv0)
{
    fprintf(fp,
            "Usage: %s \n"
            "          --thread #_of_threads_used_in_this_process\n"
            "          --next-setup #_of_threads_to_build_next_index_in_background\n"
            "          --particle-files particle_files(1 or more) [--compressed]\n"
            "          --halo-file halo_file [--filter-file filter_id_list_file]\n"
            "          --radius minimum_radius:maximum_radius:#_of_bins\n"
            "          [--backlog max_#_of_backlog_indexes]\n"
            "          [--first-setup #_of_indexes_be_firstly_built:#_of_threads_to_build_each_index]\n"
            "          [--output-file output_file_name]\n",
            argv0);
}
