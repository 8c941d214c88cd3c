/* Plain-C use of the C ABI (compiled as C99 with -Wall -Wextra -Werror by
 * tests/test_abi.py): the header must be valid C, and the host-side entry
 * points must work without a GPU.  Prints one line per check. */
#include <stdio.h>
#include <string.h>

#include "ddfan.h"

int main(void) {
    float xds[8], yds[8];
    unsigned char mask[6 * 5];
    int views[16], n = 0, bad = 0;
    if (ddfan_detector_arc(8, 4.5f, 1.2f, 1085.6f, 595.0f, xds, yds) != DDFAN_OK) bad |= 1;
    printf("arc %.6f %.6f\n", (double)xds[0], (double)yds[7]);
    if (ddfan_recon_mask(6, 5, mask) != DDFAN_OK) bad |= 2;
    int ones = 0;
    for (int i = 0; i < 30; ++i) ones += mask[i];
    printf("mask %d\n", ones);
    if (ddfan_subset_views(16, 5, 2, views, &n) != DDFAN_OK || n != 3 || views[0] != 2 ||
        views[2] != 12)
        bad |= 4;
    printf("subset %d %d %d %d\n", n, views[0], views[1], views[2]);
    if (ddfan_subset_views(16, 0, 0, views, &n) != DDFAN_EINVAL) bad |= 8;  /* numOS 0 */
    printf("einval %s\n", ddfan_last_error());
    if (strcmp(ddfan_status_string(DDFAN_EBUSY), "rejected by memory budget") != 0) bad |= 16;
    printf("version %d\n", ddfan_version());
    return bad;
}
