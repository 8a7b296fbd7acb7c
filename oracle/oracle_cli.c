/* oracle_cli.c -- TEST INFRASTRUCTURE ONLY: runs the oracle on one deck and writes its outputs.
 * usage: oracle_cli <deck> <data_dir> [max_cycles] */
#include "exaca_oracle.h"

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <deck> <data_dir> [max_cycles]\n", argv[0]);
        return 2;
    }
    oracle_sim *s;
    if (oracle_load(argv[1], argv[2], &s)) {
        fprintf(stderr, "oracle: %s\n", oracle_last_error());
        return 1;
    }
    double tcap = 0;
    int cyc = oracle_run(s, argc > 3 ? atoi(argv[3]) : 0, 1, &tcap);
    int64_t d[17];
    oracle_dims(s, d);
    printf("final cycle %d nn %lld frontier-sum %lld captures %lld solidified %lld capture-time %.3f s\n", cyc,
           (long long)d[8], (long long)d[9], (long long)d[10], (long long)d[11], tcap);
    if (oracle_write_outputs(s, argv[1])) fprintf(stderr, "oracle: %s\n", oracle_last_error());
    oracle_free(s);
    return 0;
}
