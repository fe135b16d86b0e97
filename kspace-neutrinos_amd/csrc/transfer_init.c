/* transfer_init.c -- CAMB transfer-table reader (host only).
 *
 * Same API and behaviour as transfer_init.c:9-88 of the reference: keep the
 * rows of an 8+-column CAMB transfer file with k > pi/BoxSize (in the file's
 * units), storing log k in internal units and T_nu / T_non-nu (columns 6 and
 * 8).  Errors 2019/2020 (unreadable) and 2021 (short file) are fatal.
 */
#include <math.h>
#include <stdio.h>

#include "kn_internal.h"

static int parse8(const char *line, double v[8])
{
    return sscanf(line, " %lg %lg %lg %lg %lg %lg %lg %lg", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                  &v[7]) == 8;
}

void allocate_transfer_init_table(_transfer_init_table *t, const double BoxSize, const double UnitLength_in_cm,
                                  const double InputSpectrum_UnitLength_in_cm, const char *file)
{
    const double scale = InputSpectrum_UnitLength_in_cm / UnitLength_in_cm;
    const double kmin = M_PI / BoxSize * scale;
    char line[1000];
    double v[8];
    int n = 0;
    FILE *fd = fopen(file, "r");
    if (!fd) kn_terminate(2019, "cannot open the transfer-function file %s\n", file);
    t->NPowerTable = 0;
    while (fgets(line, sizeof line, fd)) {
        if (line[0] == '#') continue;
        if (!parse8(line, v)) break;
        if (v[0] > kmin) t->NPowerTable++;
    }
    fclose(fd);
    kn_message(1, "transfer function: %d rows above k = %g\n", t->NPowerTable, kmin);
    t->logk = kn_malloc("Transfer_functions", 2 * (size_t)t->NPowerTable * sizeof(double));
    t->T_nu = t->logk + t->NPowerTable;
    fd = fopen(file, "r");
    if (!fd) kn_terminate(2020, "cannot open the transfer-function file %s\n", file);
    while (n < t->NPowerTable && fgets(line, sizeof line, fd)) {
        if (line[0] == '#') continue;
        if (!parse8(line, v)) break;
        if (v[0] > kmin) {
            t->T_nu[n] = v[5] / v[7];
            t->logk[n] = log(v[0] / scale);
            n++;
        }
    }
    fclose(fd);
    if (n < t->NPowerTable)
        kn_terminate(2021, "%d rows expected in %s, %d found\n", t->NPowerTable, file, n);
}

void free_transfer_init_table(_transfer_init_table *t) { kn_free(t->logk); }
