// sphg_solve: runs project directories the way the reference's executable does (one input.dat
// per directory, results written beside it), on the GPU.
//
//   sphg_solve [-m] <dir> [<dir> ...]      -m: mixed arithmetic form (default FP64)
#include <cstdio>
#include <cstring>

#include "../../include/sphg.h"

int main(int argc, char **argv) {
    int precision = SPHG_FP64, first = 1, result = 0;
    if (argc > 1 && strcmp(argv[1], "-m") == 0) {
        precision = SPHG_MIXED;
        first = 2;
    }
    if (first >= argc) {
        fprintf(stderr, "usage: %s [-m] <project dir> [<project dir> ...]\n", argv[0]);
        return 2;
    }
    for (int i = first; i < argc; i++) {
        int last = 0;
        const int rc = sphg_solve_project(argv[i], precision, &last);
        printf("%s: %s (last output step %d)\n", argv[i], rc ? sphg_error_string(rc) : "finished", last);
        fflush(stdout);
        if (rc && !result) result = rc;
    }
    return result ? 1 : 0;
}
