/*
 * Native environment callbacks for the "callback" I/O module
 * (io/impl/callback_module.cpp in the reference): plain C functions with the
 * io callback signature (int id, int size, void *buffer) that a caller
 * registers by address with add_io_callback (extern.h), so that per-step
 * input and output cost what the engine's copies and synchronisation cost,
 * not what a Python callback costs.
 *
 *   sg_io_normal_input     fills `size` floats with N(mean, std) values of a
 *                          counter-based stream keyed by (id, step, index);
 *                          the step counts the calls for that id
 *   sg_io_table_input      copies the next row of a table of recorded
 *                          inputs (sg_io_set_table), cycling through its
 *                          rows: a replayed stimulus, whose per-step cost is
 *                          one copy
 *   sg_io_checksum_output  folds the `size` 32-bit words of the buffer into
 *                          a running 64-bit FNV-1a checksum per id
 *
 * The ids index 64 slots; sg_io_configure sets an input slot's mean and
 * standard deviation and resets its step and checksum, sg_io_checksum reads
 * an output slot.  Not thread safe across calls for the same id (the engine
 * calls each callback from its one host thread).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SLOTS 64

static uint64_t g_step[SLOTS];
static double g_mean[SLOTS], g_std[SLOTS];
static uint64_t g_sum[SLOTS];
static float *g_table[SLOTS];
static int g_rows[SLOTS], g_cols[SLOTS];

/* splitmix64 (Steele, Lea, Flood, OOPSLA 2014) of a counter */
static uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void sg_io_configure(int id, double mean, double std) {
    if (id < 0 || id >= SLOTS) return;
    g_mean[id] = mean;
    g_std[id] = std;
    g_step[id] = 0;
    g_sum[id] = 0xCBF29CE484222325ull;   /* FNV-1a offset basis */
}

/* Keeps a copy of `rows` x `cols` floats for sg_io_table_input on slot
 * `id` and resets its step; returns 0, or -1 (bad id or size, no memory). */
int sg_io_set_table(int id, const float *table, int rows, int cols) {
    if (id < 0 || id >= SLOTS || rows < 1 || cols < 1) return -1;
    float *t = (float *)malloc((size_t)rows * cols * sizeof(float));
    if (!t) return -1;
    memcpy(t, table, (size_t)rows * cols * sizeof(float));
    free(g_table[id]);
    g_table[id] = t;
    g_rows[id] = rows;
    g_cols[id] = cols;
    g_step[id] = 0;
    return 0;
}

void sg_io_table_input(int id, int size, void *buffer) {
    if (id < 0 || id >= SLOTS || !g_table[id]) return;
    const int row = (int)(g_step[id]++ % (uint64_t)g_rows[id]);
    const int n = size < g_cols[id] ? size : g_cols[id];
    memcpy(buffer, g_table[id] + (size_t)row * g_cols[id],
           (size_t)n * sizeof(float));
}

uint64_t sg_io_checksum(int id) {
    return (id >= 0 && id < SLOTS) ? g_sum[id] : 0;
}

void sg_io_normal_input(int id, int size, void *buffer) {
    if (id < 0 || id >= SLOTS) return;
    float *out = (float *)buffer;
    const uint64_t base = mix(((uint64_t)id << 48) ^ g_step[id]++);
    for (int i = 0; i < size; ++i) {
        const uint64_t r = mix(base + (uint64_t)i);
        /* Box-Muller over two 26-bit uniforms in (0, 1] and [0, 1) */
        const double u1 = ((double)(r >> 38) + 1.0) / 67108864.0;
        const double u2 = (double)((r >> 12) & 0x3FFFFFFull) / 67108864.0;
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        out[i] = (float)(g_mean[id] + g_std[id] * z);
    }
}

void sg_io_checksum_output(int id, int size, void *buffer) {
    if (id < 0 || id >= SLOTS) return;
    const uint32_t *w = (const uint32_t *)buffer;
    uint64_t h = g_sum[id];
    for (int i = 0; i < size; ++i) h = (h ^ w[i]) * 0x100000001B3ull;
    g_sum[id] = h;
}
