/*
 * swe_nc2bin: the reference's nc -> bin converter (swe_code/io/nc2bin.c)
 * on this repository's NetCDF reader.  Same parameters: the environment
 * variables NC2BIN_INPUT_FILE (NetCDF classic file) and NC2BIN_OUTPUT_FILE
 * (binary input file written).  Exit 0 on success, 2 if the input cannot be
 * opened, 1 on any other error.
 */
#include "swe_host.h"

#include <stdio.h>
#include <stdlib.h>

int main(void)
{
    const char *in = getenv("NC2BIN_INPUT_FILE");
    const char *out = getenv("NC2BIN_OUTPUT_FILE");
    if (!in || !out) {
        fprintf(stderr, "swe_nc2bin: set NC2BIN_INPUT_FILE and NC2BIN_OUTPUT_FILE\n");
        return 1;
    }
    swe_gsmd g;
    int err = swe_get_gsmd_nc(in, &g);
    if (!err) {
        printf("\nNC2BIN:\tNnodes = %d,\tNnbr = %d\n", g.Nnodes, g.Nnbr);
        swe_gsmd_free(&g);
        err = swe_nc2bin(in, out);
    }
    if (err == SWE_HOST_EOPEN) {
        printf("Error in input.c: Could not open %s \n", in);
        return 2;
    }
    if (err) {
        fprintf(stderr, "swe_nc2bin: %s: host error %d\n", in, err);
        return 1;
    }
    return 0;
}
