// Host check of the FP16C storage format (d3q19_common.h: fp16c_pack,
// fp16c_unpack; 1 sign, 4 exponent, 11 significand bits of values scaled by
// 2^-113), against what the format's definition fixes:
//  * every one of the 65536 words decodes to the value
//    (-1)^s 2^(e-14) (1 + m/2048) for e >= 1, (-1)^s 2^-13 (m/2048) for
//    e = 0, and encodes back to itself;
//  * positive words decode in increasing order;
//  * encoding a float gives the nearest word (found by brute force over all
//    words, ties to an even significand) below the largest value, and the
//    largest value above it, and for NaN, with the sign kept.
// Built with hipcc and run on the CPU; launches nothing on a GPU.
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <random>

#include "d3q19_common.h"

using namespace d3q19;

static double word_value(uint32_t h)
{
    const int s = (h >> 15) & 1, e = (h >> 11) & 15, m = h & 0x7ff;
    const double v = e ? ldexp(1.0 + m / 2048.0, e - 14) : ldexp(m / 2048.0, -13);
    return s ? -v : v;
}

static double decode(uint16_t h) { return (double)fp16c_unpack(h) * (double)FP16C_UP; }

static uint16_t encode(double x) { return fp16c_pack((float)x * FP16C_DOWN); }

int main()
{
    // 1. decoding and round trip of every word
    for (uint32_t h = 0; h < 65536; h++) {
        const double want = word_value(h);
        if (decode((uint16_t)h) != want) {
            fprintf(stderr, "word %04x decodes to %.17g, want %.17g\n", h, decode((uint16_t)h), want);
            return 1;
        }
        if (encode(want) != (uint16_t)h) {
            fprintf(stderr, "word %04x: value %.17g encodes to %04x\n", h, want, encode(want));
            return 1;
        }
    }
    // 2. order
    for (uint32_t h = 0; h + 1 < 0x8000; h++)
        if (!(decode((uint16_t)(h + 1)) > decode((uint16_t)h))) {
            fprintf(stderr, "words %04x and %04x out of order\n", h, h + 1);
            return 1;
        }
    // named values
    if (decode(0x7fff) != 4.0 - ldexp(1.0, -10) || decode(0x0800) != ldexp(1.0, -13) ||
        decode(0x0001) != ldexp(1.0, -24)) {
        fprintf(stderr, "largest / smallest normal / smallest denormal wrong\n");
        return 1;
    }
    // 3. nearest word, brute force, for values across the whole range (FP32
    // inputs: the pack rounds the FP32 value)
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ex(-30.0, 3.0);
    long long checked = 0;
    for (int t = 0; t < 4000; t++) {
        float x = (float)ldexp(1.0 + (rng() >> 11) * 0x1.0p-53, (int)floor(ex(rng)));
        if (t % 3 == 0)
            x = -x;
        if (t % 50 == 0) // an exact tie between two neighbouring words
            x = (float)(0.5 * (word_value(0x1234 + t) + word_value(0x1235 + t)));
        const double ax = fabs((double)x);
        uint32_t best = 0;
        double bd = INFINITY;
        for (uint32_t h = 0; h < 0x8000; h++) {
            const double d = fabs(word_value(h) - ax);
            if (d < bd || (d == bd && (h & 1) == 0 && (best & 1) == 1)) {
                bd = d;
                best = h;
            }
        }
        if (ax > word_value(0x7fff))
            best = 0x7fff;
        if (x < 0)
            best |= 0x8000;
        const uint16_t got = encode(x);
        if (got != best) {
            fprintf(stderr, "x = %.9g encodes to %04x (%.17g), nearest %04x (%.17g)\n", (double)x, got, decode(got),
                    best, word_value(best));
            return 1;
        }
        checked++;
    }
    // saturation and NaN
    if (encode(5.0) != 0x7fff || encode(-1e30) != 0xffff || (encode(NAN) & 0x7fff) != 0x7fff) {
        fprintf(stderr, "saturation wrong\n");
        return 1;
    }
    printf("ok fp16c words 65536 nearest %lld\n", checked);
    return 0;
}
