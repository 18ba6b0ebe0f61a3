/*
 * mkl_streams.c -- feeds the oracle's rank mode with the reference's own
 * noise: Intel MKL VSL MT19937 streams, Gaussians by ICDF and uniforms by
 * UNIFORM_STD_ACCURATE (RANDOMNUMBER.cpp:84-91, 175-182, 268-274).
 *
 * TEST INFRASTRUCTURE ONLY, and only where an MKL runtime exists (this
 * build container's conda libmkl_rt).  MKL is loaded with dlopen so that
 * nothing here needs MKL headers; the library is called, never inspected.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>

typedef int (*new_stream_fn)(void** stream, int brng, unsigned int seed);
typedef int (*del_stream_fn)(void** stream);
typedef int (*gauss_fn)(int method, void* stream, int n, double* r, double a, double sigma);
typedef int (*unif_fn)(int method, void* stream, int n, double* r, double a, double b);

#define VSL_BRNG_MT19937 0x800000
#define VSL_RNG_METHOD_GAUSSIAN_ICDF 2
#define VSL_RNG_METHOD_UNIFORM_STD_ACCURATE 0x40000000

static void* g_lib;
static new_stream_fn p_new;
static del_stream_fn p_del;
static gauss_fn p_gauss;
static unif_fn p_unif;

typedef struct ldm_ctx {
    void* stream[3];
} ldm_ctx;

int ldm_available(const char* path)
{
    if (g_lib) return 1;
    g_lib = dlopen(path ? path : "libmkl_rt.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!g_lib) return 0;
    p_new = (new_stream_fn)dlsym(g_lib, "vslNewStream");
    p_del = (del_stream_fn)dlsym(g_lib, "vslDeleteStream");
    p_gauss = (gauss_fn)dlsym(g_lib, "vdRngGaussian");
    p_unif = (unif_fn)dlsym(g_lib, "vdRngUniform");
    return p_new && p_del && p_gauss && p_unif;
}

/* seeds for the Gaussian, SIMD-uniform and scalar-uniform streams */
ldm_ctx* ldm_create(unsigned int seed_g, unsigned int seed_u, unsigned int seed_us)
{
    ldm_ctx* c = (ldm_ctx*)calloc(1, sizeof(ldm_ctx));
    unsigned int seeds[3] = {seed_g, seed_u, seed_us};
    for (int k = 0; k < 3; ++k) {
        if (p_new(&c->stream[k], VSL_BRNG_MT19937, seeds[k]) != 0) {
            free(c);
            return NULL;
        }
    }
    return c;
}

void ldm_destroy(ldm_ctx* c)
{
    if (!c) return;
    for (int k = 0; k < 3; ++k) p_del(&c->stream[k]);
    free(c);
}

/* ldo_fill_fn */
void ldm_fill(void* ctx, int which, double* buf, uint64_t n)
{
    ldm_ctx* c = (ldm_ctx*)ctx;
    if (which == 0) p_gauss(VSL_RNG_METHOD_GAUSSIAN_ICDF, c->stream[0], (int)n, buf, 0.0, 1.0);
    else p_unif(VSL_RNG_METHOD_UNIFORM_STD_ACCURATE, c->stream[which], (int)n, buf, 0.0, 1.0);
}

void* ldm_fill_ptr(void) { return (void*)&ldm_fill; }
