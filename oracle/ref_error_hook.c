/* Error hook linked into oracle/_ref/libquest_ref.so (QuEST's own CPU backend).
 *
 * QuEST declares invalidQuESTInputError weak (QuEST_validation.c:240-245) so a
 * caller can replace its exit(); this strong definition records the message
 * before exiting, so a harness that reads stderr sees which call was invalid.
 * Test infrastructure only: the harnesses never pass invalid input to it. */
#include <stdio.h>
#include <stdlib.h>

void invalidQuESTInputError(const char *errMsg, const char *errFunc) {
    fprintf(stderr, "QuEST reference error in %s: %s\n", errFunc, errMsg);
    fflush(stderr);
    exit(EXIT_FAILURE);
}
