// Writes Philox4x32-10 output words of rocRAND's engine (seed k, subsequence
// 0, offset 0: counter (b,0,0,0), key (lo32 k, hi32 k)) as golden vectors for
// the oracle's stream.  Host-only use of rocRAND's header implementation.
#include <cstdio>
#include <rocrand/rocrand_kernel.h>

int main()
{
    const unsigned long long seeds[] = {0ull, 1ull, 7ull, 4095ull, 0x123456789abcdefull,
                                        0xffffffffffffffffull};
    printf("{\n  \"vectors\": [\n");
    for (int s = 0; s < 6; ++s) {
        rocrand_device::philox4x32_10_engine e(seeds[s], 0, 0);
        printf("    {\"seed\": \"%llu\", \"u32\": [", seeds[s]);
        for (int i = 0; i < 16; ++i)
            printf("%s%u", i ? ", " : "", e.next());
        printf("]}%s\n", s < 5 ? "," : "");
    }
    printf("  ]\n}\n");
    return 0;
}
