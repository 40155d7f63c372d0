/* ringcheck.c - test infrastructure: whether the members of a Moore ring
 * are one piece, by two methods, over a range of masks.
 *
 * connectivity.rs:43-84 (locally_connected) asks whether the old label's
 * sites among the target's ring form one piece, two positions joined when
 * they are within one step on every axis (connectivity.rs:31-35). The
 * emulation answers by a breadth-first search over the mask's bits
 * (emu_one_piece); this file answers by union-find over an adjacency list
 * built from the ring's coordinates, so the two can be compared on every
 * mask of the planar ring (2^8) and of the volume's (2^26). */
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

int emu_one_piece(uint32_t present, int ring_size);

static int find(int *parent, int a) {
    while (parent[a] != a) {
        parent[a] = parent[parent[a]];
        a = parent[a];
    }
    return a;
}

/* Ring positions in connectivity.rs:23-29's order: dz, dy, dx from -1 to
 * 1, centre dropped, z offsets dropped in a plane. */
static int ring_positions(int ring_size, int pos[26][3]) {
    int k = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                if (ring_size == 8 && dz) continue;
                pos[k][0] = dx;
                pos[k][1] = dy;
                pos[k][2] = dz;
                ++k;
            }
    return k;
}

static int uf_one_piece(uint32_t mask, int n, const int adj[26][26], int *degree) {
    int parent[26], members = 0, root = -1;
    for (int a = 0; a < n; ++a) parent[a] = a;
    for (int a = 0; a < n; ++a) {
        if (!(mask >> a & 1u)) continue;
        ++members;
        for (int j = 0; j < degree[a]; ++j) {
            int b = adj[a][j];
            if (!(mask >> b & 1u)) continue;
            int ra = find(parent, a), rb = find(parent, b);
            if (ra != rb) parent[ra] = rb;
        }
    }
    if (members <= 1) return 1;
    for (int a = 0; a < n; ++a) {
        if (!(mask >> a & 1u)) continue;
        int r = find(parent, a);
        if (root < 0) root = r;
        else if (r != root) return 0;
    }
    return 1;
}

/* One byte per mask in [lo, hi): bit 0 the union-find answer, bit 1 the
 * emulation's. */
void oracle_ring_table(int ring_size, uint32_t lo, uint32_t hi, uint8_t *out) {
    int pos[26][3], adj[26][26], degree[26];
    int n = ring_positions(ring_size, pos);
    for (int a = 0; a < n; ++a) {
        degree[a] = 0;
        for (int b = 0; b < n; ++b) {
            if (a == b) continue;
            if (abs(pos[a][0] - pos[b][0]) <= 1 && abs(pos[a][1] - pos[b][1]) <= 1 &&
                abs(pos[a][2] - pos[b][2]) <= 1)
                adj[a][degree[a]++] = b;
        }
    }
    for (uint32_t m = lo; m < hi; ++m)
        out[m - lo] = (uint8_t)(uf_one_piece(m, n, adj, degree) | (emu_one_piece(m, ring_size) << 1));
}
