/*
 * Driver for the reference rannew64 (rannew64.cu:55-126), compiled from the
 * reference source where it lies (see oracle/Makefile, target `ref`).
 * TEST INFRASTRUCTURE ONLY: writes golden vectors, one seed per process,
 * because the reference keeps its generator state in statics.
 *
 * usage: rannew64_ref IDUM COUNT   -> COUNT lines, each a C99 hex double
 */
#include <stdio.h>
#include <stdlib.h>

double rannew64(long* pidum);

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IDUM COUNT\n", argv[0]);
    return 2;
  }
  long idum = strtol(argv[1], NULL, 10);
  long count = strtol(argv[2], NULL, 10);
  for (long i = 0; i < count; i++) printf("%a\n", rannew64(&idum));
  return 0;
}
