/*
 * Oracle harness: runs the reference's own GLSL compute program under Mesa
 * llvmpipe (OpenGL 4.3+ compute) with no window system.
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ from the reference's
 * sources (oracle/Makefile), used by tests/golden/make_golden.py to write
 * fixtures and by bench.py's cpu_baseline leg as the timed CPU baseline.
 *
 * What it restates from the reference:
 *   build_compute_shaders()  src/lib.rs:31-75  (textual #include resolver)
 *   Simulation::new          sandengine-core/src/simulation.rs:128-192
 *                            (RGBA32F textures, 256-entry modification UBO,
 *                            dispatch geometry ((W+7)/8, (H+7)/8, 1))
 *   Simulation::run          sandengine-core/src/simulation.rs:195-252
 *                            (frame += 1, copy mods, dispatch, ping-pong swap,
 *                            zero the used mod_size slots)
 * The context comes from Mesa's swrast DRI driver through the public
 * GL/internal/dri_interface.h API; GL entry points come from libglapi.
 */
#define GL_GLEXT_PROTOTYPES 0
#include <GL/glcorearb.h>
#include <GL/internal/dri_interface.h>
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

extern const char se_glsl_falling_sand[], se_glsl_operations[], se_glsl_math[], se_glsl_cell[],
    se_glsl_directions[], se_glsl_material_helpers[], se_glsl_gen_materials[], se_glsl_gen_rules[],
    se_glsl_gen_falling_sand[];

static char g_err[8192];
static void seterr(const char *s) { snprintf(g_err, sizeof g_err, "%s", s); }
const char *lp_error(void) { return g_err; }

/* ------------------------------------------------------------------ GL loading */
#define GLFN(T, n) static T n##_p;
#define GL_LIST(X) \
    X(PFNGLGETSTRINGPROC, glGetString) X(PFNGLGETINTEGERVPROC, glGetIntegerv) \
    X(PFNGLGETERRORPROC, glGetError) X(PFNGLGENTEXTURESPROC, glGenTextures) \
    X(PFNGLDELETETEXTURESPROC, glDeleteTextures) X(PFNGLBINDTEXTUREPROC, glBindTexture) \
    X(PFNGLTEXIMAGE2DPROC, glTexImage2D) X(PFNGLTEXSUBIMAGE2DPROC, glTexSubImage2D) \
    X(PFNGLTEXPARAMETERIPROC, glTexParameteri) X(PFNGLGETTEXIMAGEPROC, glGetTexImage) \
    X(PFNGLACTIVETEXTUREPROC, glActiveTexture) X(PFNGLCREATESHADERPROC, glCreateShader) \
    X(PFNGLSHADERSOURCEPROC, glShaderSource) X(PFNGLCOMPILESHADERPROC, glCompileShader) \
    X(PFNGLGETSHADERIVPROC, glGetShaderiv) X(PFNGLGETSHADERINFOLOGPROC, glGetShaderInfoLog) \
    X(PFNGLCREATEPROGRAMPROC, glCreateProgram) X(PFNGLATTACHSHADERPROC, glAttachShader) \
    X(PFNGLLINKPROGRAMPROC, glLinkProgram) X(PFNGLGETPROGRAMIVPROC, glGetProgramiv) \
    X(PFNGLGETPROGRAMINFOLOGPROC, glGetProgramInfoLog) X(PFNGLUSEPROGRAMPROC, glUseProgram) \
    X(PFNGLDELETEPROGRAMPROC, glDeleteProgram) X(PFNGLDELETESHADERPROC, glDeleteShader) \
    X(PFNGLGETUNIFORMLOCATIONPROC, glGetUniformLocation) X(PFNGLUNIFORM1IPROC, glUniform1i) \
    X(PFNGLUNIFORM2IPROC, glUniform2i) X(PFNGLUNIFORM1FPROC, glUniform1f) \
    X(PFNGLGETUNIFORMBLOCKINDEXPROC, glGetUniformBlockIndex) \
    X(PFNGLUNIFORMBLOCKBINDINGPROC, glUniformBlockBinding) \
    X(PFNGLGETACTIVEUNIFORMBLOCKIVPROC, glGetActiveUniformBlockiv) \
    X(PFNGLGETUNIFORMINDICESPROC, glGetUniformIndices) \
    X(PFNGLGETACTIVEUNIFORMSIVPROC, glGetActiveUniformsiv) X(PFNGLGENBUFFERSPROC, glGenBuffers) \
    X(PFNGLBINDBUFFERPROC, glBindBuffer) X(PFNGLBUFFERDATAPROC, glBufferData) \
    X(PFNGLBUFFERSUBDATAPROC, glBufferSubData) X(PFNGLBINDBUFFERBASEPROC, glBindBufferBase) \
    X(PFNGLBINDIMAGETEXTUREPROC, glBindImageTexture) \
    X(PFNGLDISPATCHCOMPUTEPROC, glDispatchCompute) X(PFNGLMEMORYBARRIERPROC, glMemoryBarrier) \
    X(PFNGLFINISHPROC, glFinish) X(PFNGLPIXELSTOREIPROC, glPixelStorei)
GL_LIST(GLFN)

static void gdi(__DRIdrawable *d, int *x, int *y, int *w, int *h, void *p) {
    (void)d; (void)p; *x = *y = 0; *w = *h = 1;
}
static void pimg(__DRIdrawable *d, int op, int x, int y, int w, int h, char *data, void *p) {
    (void)d; (void)op; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static void gimg(__DRIdrawable *d, int x, int y, int w, int h, char *data, void *p) {
    (void)d; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static const __DRIswrastLoaderExtension g_swl = {{__DRI_SWRAST_LOADER, 1}, gdi, pimg, gimg};
static const __DRIextension *g_loader_exts[] = {&g_swl.base, NULL};

static int g_ready = 0;
static char g_info[512];

int lp_init(int nthreads) {
    if (g_ready) return 0;
    if (nthreads > 0) {
        char b[32];
        snprintf(b, sizeof b, "%d", nthreads);
        setenv("LP_NUM_THREADS", b, 1);
    }
    setenv("GALLIUM_DRIVER", "llvmpipe", 1);
    void *glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!glapi) { seterr(dlerror()); return -1; }
    const char *paths[] = {"/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so",
                           "/usr/lib64/dri/swrast_dri.so", "/usr/lib/dri/swrast_dri.so", NULL};
    void *drv = NULL;
    for (int i = 0; paths[i] && !drv; i++) drv = dlopen(paths[i], RTLD_NOW | RTLD_GLOBAL);
    if (!drv) { seterr("swrast_dri.so not found"); return -1; }
    typedef const __DRIextension **(*getext_t)(void);
    getext_t ge = (getext_t)dlsym(drv, "__driDriverGetExtensions_swrast");
    if (!ge) { seterr("__driDriverGetExtensions_swrast missing"); return -1; }
    const __DRIextension **ext = ge();
    const __DRIcoreExtension *core = NULL;
    const __DRIswrastExtension *sw = NULL;
    for (int i = 0; ext[i]; i++) {
        if (!strcmp(ext[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)ext[i];
        if (!strcmp(ext[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension *)ext[i];
    }
    if (!core || !sw || sw->base.version < 4) { seterr("DRI core/swrast v4 missing"); return -1; }
    const __DRIconfig **cfgs = NULL;
    __DRIscreen *scr = sw->createNewScreen2(0, g_loader_exts, ext, &cfgs, NULL);
    if (!scr) { seterr("createNewScreen2 failed"); return -1; }
    uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
    unsigned err = 0;
    __DRIcontext *ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE, cfgs ? cfgs[0] : NULL,
                                                 NULL, 2, attribs, &err, NULL);
    if (!ctx) { snprintf(g_err, sizeof g_err, "createContextAttribs failed (%u)", err); return -1; }
    if (!core->bindContext(ctx, NULL, NULL)) { seterr("bindContext failed"); return -1; }
    typedef void *(*gpa_t)(const char *);
    gpa_t gpa = (gpa_t)dlsym(glapi, "_glapi_get_proc_address");
    if (!gpa) { seterr("_glapi_get_proc_address missing"); return -1; }
#define GLLOAD(T, n) n##_p = (T)gpa(#n); if (!n##_p) { seterr("missing " #n); return -1; }
    GL_LIST(GLLOAD)
    GLint mts = 0;
    glGetIntegerv_p(GL_MAX_TEXTURE_SIZE, &mts);
    snprintf(g_info, sizeof g_info, "%s | %s | max_texture %d | LP_NUM_THREADS=%s",
             (const char *)glGetString_p(GL_VERSION), (const char *)glGetString_p(GL_RENDERER), mts,
             getenv("LP_NUM_THREADS") ? getenv("LP_NUM_THREADS") : "default");
    g_ready = 1;
    return 0;
}
const char *lp_info(void) { return g_info; }

/* ------------------------------------------------------------------ include resolver */
static const char *lookup(const char *path, const char *gen_mat, const char *gen_rules) {
    if (!strcmp(path, "directions.glsl")) return se_glsl_directions;
    if (!strcmp(path, "math.glsl")) return se_glsl_math;
    if (!strcmp(path, "cell.glsl")) return se_glsl_cell;
    if (!strcmp(path, "operations.glsl")) return se_glsl_operations;
    if (!strcmp(path, "material_helpers.glsl")) return se_glsl_material_helpers;
    if (!strcmp(path, "gen/materials.glsl")) return gen_mat;
    if (!strcmp(path, "gen/rules.glsl")) return gen_rules;
    return NULL;
}

/* src/lib.rs:44-69 for one file: returns malloc'd text or NULL. */
static char *resolve(const char *src, const char *gen_mat, const char *gen_rules) {
    static const char *S = "\n#include \"";
    size_t sl = strlen(S);
    size_t cap = strlen(src) + 1;
    char *c = (char *)malloc(cap);
    memcpy(c, src, cap);
    for (int guard = 0; guard < 1000; guard++) {
        char *hit = strstr(c, S);
        if (!hit || hit == c) return c;  /* unwrap_or(0): index 0 counts as "none" */
        size_t start = (size_t)(hit - c);
        const char *p0 = hit + sl;
        const char *q = strchr(p0, '"');
        size_t plen = q ? (size_t)(q - p0) : strlen(p0);
        char path[256];
        if (plen >= sizeof path) { free(c); return NULL; }
        memcpy(path, p0, plen);
        path[plen] = 0;
        const char *inc = lookup(path, gen_mat, gen_rules);
        if (!inc) { snprintf(g_err, sizeof g_err, "include %s not found", path); free(c); return NULL; }
        size_t end = start + sl + plen + 1;
        char before = c[start - 1];
        size_t ilen = (before == ' ' || before == '/') ? 0 : strlen(inc) + 3;
        size_t tail = strlen(c + end);
        char *n = (char *)malloc(start + ilen + tail + 1);
        memcpy(n, c, start);
        if (ilen) {
            memcpy(n + start, inc, ilen - 3);
            memcpy(n + start + ilen - 3, "\n\n\n", 3);
        }
        memcpy(n + start + ilen, c + end, tail + 1);
        free(c);
        c = n;
    }
    free(c);
    return NULL;
}

/* Assembles falling_sand.glsl with the given generated parts (NULL -> the
 * reference's checked-in gen/ files).  Caller frees with lp_free. */
char *lp_assemble(const char *gen_mat, const char *gen_rules) {
    return resolve(se_glsl_falling_sand, gen_mat ? gen_mat : se_glsl_gen_materials,
                   gen_rules ? gen_rules : se_glsl_gen_rules);
}
void lp_free(void *p) { free(p); }
const char *lp_ref_gen(int which) {
    return which == 0 ? se_glsl_gen_materials : which == 1 ? se_glsl_gen_rules : se_glsl_gen_falling_sand;
}

/* ------------------------------------------------------------------ simulation */
typedef struct { int32_t position[2]; int32_t shape, size, matID, pad[3]; } lp_mod;

static GLuint g_prog, g_tex_in, g_tex_out, g_tex_color, g_light_in, g_light_out, g_ubo;
static int g_W, g_H, g_frame;
static GLint g_mod_stride = 32, g_mod_off_pos = 0, g_mod_off_shape = 8, g_mod_off_size = 12,
             g_mod_off_mat = 16;
static unsigned char *g_ubo_shadow;
static GLint g_ubo_size;
static int g_prev_mods;

int lp_program(const char *src) {
    if (!g_ready) { seterr("lp_init not called"); return -1; }
    GLuint sh = glCreateShader_p(GL_COMPUTE_SHADER);
    glShaderSource_p(sh, 1, &src, NULL);
    glCompileShader_p(sh);
    GLint ok = 0;
    glGetShaderiv_p(sh, GL_COMPILE_STATUS, &ok);
    if (!ok) { glGetShaderInfoLog_p(sh, sizeof g_err, NULL, g_err); glDeleteShader_p(sh); return -2; }
    GLuint pr = glCreateProgram_p();
    glAttachShader_p(pr, sh);
    glLinkProgram_p(pr);
    glGetProgramiv_p(pr, GL_LINK_STATUS, &ok);
    glDeleteShader_p(sh);
    if (!ok) { glGetProgramInfoLog_p(pr, sizeof g_err, NULL, g_err); glDeleteProgram_p(pr); return -3; }
    if (g_prog) glDeleteProgram_p(g_prog);
    g_prog = pr;
    /* modification UBO layout (shared layout: query the offsets) */
    GLuint bi = glGetUniformBlockIndex_p(pr, "SimModifications");
    if (bi == GL_INVALID_INDEX) { seterr("SimModifications block missing"); return -4; }
    glGetActiveUniformBlockiv_p(pr, bi, GL_UNIFORM_BLOCK_DATA_SIZE, &g_ubo_size);
    const char *names[5] = {"sim_modifications[0].position", "sim_modifications[0].mod_shape",
                            "sim_modifications[0].mod_size", "sim_modifications[0].mod_matID",
                            "sim_modifications[1].position"};
    GLuint idx[5];
    glGetUniformIndices_p(pr, 5, names, idx);
    GLint off[5] = {0, 8, 12, 16, 32};
    for (int i = 0; i < 5; i++)
        if (idx[i] != GL_INVALID_INDEX) glGetActiveUniformsiv_p(pr, 1, &idx[i], GL_UNIFORM_OFFSET, &off[i]);
    g_mod_off_pos = off[0]; g_mod_off_shape = off[1]; g_mod_off_size = off[2];
    g_mod_off_mat = off[3]; g_mod_stride = off[4] - off[0];
    glUniformBlockBinding_p(pr, bi, 0);
    return 0;
}

static GLuint mk_tex(int W, int H, const float *data) {
    GLuint t;
    glGenTextures_p(1, &t);
    glBindTexture_p(GL_TEXTURE_2D, t);
    glTexImage2D_p(GL_TEXTURE_2D, 0, GL_RGBA32F, W, H, 0, GL_RGBA, GL_FLOAT, data);
    glTexParameteri_p(GL_TEXTURE_2D, GL_TEXTURE_MIN_FILTER, GL_NEAREST);
    glTexParameteri_p(GL_TEXTURE_2D, GL_TEXTURE_MAG_FILTER, GL_NEAREST);
    glTexParameteri_p(GL_TEXTURE_2D, GL_TEXTURE_BASE_LEVEL, 0);
    glTexParameteri_p(GL_TEXTURE_2D, GL_TEXTURE_MAX_LEVEL, 0);
    return t;
}

int lp_world(int W, int H) {
    if (!g_prog) { seterr("no program"); return -1; }
    if (g_tex_in) {
        GLuint ts[5] = {g_tex_in, g_tex_out, g_tex_color, g_light_in, g_light_out};
        glDeleteTextures_p(5, ts);
    }
    g_W = W; g_H = H; g_frame = 0;
    float *z = (float *)calloc((size_t)W * H * 4, sizeof(float));
    if (!z) { seterr("oom"); return -1; }
    glPixelStorei_p(GL_UNPACK_ALIGNMENT, 4);
    glPixelStorei_p(GL_PACK_ALIGNMENT, 4);
    g_tex_in = mk_tex(W, H, z); g_tex_out = mk_tex(W, H, z); g_tex_color = mk_tex(W, H, z);
    g_light_in = mk_tex(W, H, z); g_light_out = mk_tex(W, H, z);
    free(z);
    if (!g_ubo) glGenBuffers_p(1, &g_ubo);
    free(g_ubo_shadow);
    g_ubo_shadow = (unsigned char *)calloc(1, (size_t)g_ubo_size);
    glBindBuffer_p(GL_UNIFORM_BUFFER, g_ubo);
    glBufferData_p(GL_UNIFORM_BUFFER, g_ubo_size, g_ubo_shadow, GL_DYNAMIC_DRAW);
    g_prev_mods = 0;
    return glGetError_p() == GL_NO_ERROR ? 0 : -2;
}

int lp_upload(const uint32_t *grid) {
    size_t n = (size_t)g_W * g_H;
    float *f = (float *)calloc(n * 4, sizeof(float));
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) f[4 * i] = (float)grid[i];
    glBindTexture_p(GL_TEXTURE_2D, g_tex_in);
    glTexSubImage2D_p(GL_TEXTURE_2D, 0, 0, 0, g_W, g_H, GL_RGBA, GL_FLOAT, f);
    free(f);
    return glGetError_p() == GL_NO_ERROR ? 0 : -2;
}

int lp_download(uint32_t *grid) {
    size_t n = (size_t)g_W * g_H;
    float *f = (float *)malloc(n * 4 * sizeof(float));
    if (!f) return -1;
    glFinish_p();
    glBindTexture_p(GL_TEXTURE_2D, g_tex_in);
    glGetTexImage_p(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, f);
    for (size_t i = 0; i < n; i++) grid[i] = (uint32_t)(int32_t)f[4 * i];
    free(f);
    return glGetError_p() == GL_NO_ERROR ? 0 : -2;
}

/* RGBA f32 images of the last step: colour (output_color) and light (the
 * light that step wrote, held in the light input after the swap). */
static int download_rgba(GLuint tex, float *out) {
    glFinish_p();
    glBindTexture_p(GL_TEXTURE_2D, tex);
    glGetTexImage_p(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, out);
    return glGetError_p() == GL_NO_ERROR ? 0 : -2;
}
int lp_download_color(float *rgba) { return download_rgba(g_tex_color, rgba); }
int lp_download_light(float *rgba) { return download_rgba(g_light_in, rgba); }

int lp_get_frame(void) { return g_frame; }
void lp_set_frame(int f) { g_frame = f; }

static void put_mods(const lp_mod *mods, int n) {
    for (int i = 0; i < n && i < 256; i++) {
        unsigned char *b = g_ubo_shadow + (size_t)i * g_mod_stride;
        memcpy(b + g_mod_off_pos, mods[i].position, 8);
        memcpy(b + g_mod_off_shape, &mods[i].shape, 4);
        memcpy(b + g_mod_off_size, &mods[i].size, 4);
        memcpy(b + g_mod_off_mat, &mods[i].matID, 4);
    }
}

/* One Simulation::run(): frame += 1, upload mods, dispatch, swap, zero sizes. */
static void one_step(const lp_mod *mods, int nmods) {
    g_frame += 1;
    if (nmods > 256) nmods = 256;
    if (nmods > 0) {
        put_mods(mods, nmods);
        glBindBuffer_p(GL_UNIFORM_BUFFER, g_ubo);
        glBufferSubData_p(GL_UNIFORM_BUFFER, 0, g_ubo_size, g_ubo_shadow);
    }
    glUseProgram_p(g_prog);
    glActiveTexture_p(GL_TEXTURE0);
    glBindTexture_p(GL_TEXTURE_2D, g_tex_in);
    glActiveTexture_p(GL_TEXTURE4);
    glBindTexture_p(GL_TEXTURE_2D, g_light_in);
    GLint l;
    if ((l = glGetUniformLocation_p(g_prog, "input_data")) >= 0) glUniform1i_p(l, 0);
    if ((l = glGetUniformLocation_p(g_prog, "input_light")) >= 0) glUniform1i_p(l, 4);
    glBindImageTexture_p(0, g_tex_out, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_RGBA32F);
    glBindImageTexture_p(1, g_tex_color, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_RGBA32F);
    glBindImageTexture_p(5, g_light_out, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_RGBA32F);
    if ((l = glGetUniformLocation_p(g_prog, "output_data")) >= 0) glUniform1i_p(l, 0);
    if ((l = glGetUniformLocation_p(g_prog, "output_color")) >= 0) glUniform1i_p(l, 1);
    if ((l = glGetUniformLocation_p(g_prog, "simSize")) >= 0) glUniform2i_p(l, g_W, g_H);
    if ((l = glGetUniformLocation_p(g_prog, "frame")) >= 0) glUniform1i_p(l, g_frame);
    if ((l = glGetUniformLocation_p(g_prog, "moveRight")) >= 0) glUniform1i_p(l, 1);
    if ((l = glGetUniformLocation_p(g_prog, "time")) >= 0) glUniform1f_p(l, 0.0f);
    glBindBufferBase_p(GL_UNIFORM_BUFFER, 0, g_ubo);
    glDispatchCompute_p((GLuint)((g_W + 7) / 8), (GLuint)((g_H + 7) / 8), 1);
    glMemoryBarrier_p(GL_ALL_BARRIER_BITS);
    GLuint t = g_tex_in; g_tex_in = g_tex_out; g_tex_out = t;
    t = g_light_in; g_light_in = g_light_out; g_light_out = t;
    if (nmods > 0) {
        for (int i = 0; i < nmods; i++) {
            int32_t z = 0;
            memcpy(g_ubo_shadow + (size_t)i * g_mod_stride + g_mod_off_size, &z, 4);
        }
        glBindBuffer_p(GL_UNIFORM_BUFFER, g_ubo);
        glBufferSubData_p(GL_UNIFORM_BUFFER, 0, g_ubo_size, g_ubo_shadow);
    }
}

/* nsteps runs of Simulation::run; mods (n x 32 B) apply to the first. */
int lp_run(int nsteps, const lp_mod *mods, int nmods) {
    for (int s = 0; s < nsteps; s++) one_step(s == 0 ? mods : NULL, s == 0 ? nmods : 0);
    glFinish_p();
    GLenum e = glGetError_p();
    if (e != GL_NO_ERROR) { snprintf(g_err, sizeof g_err, "GL error 0x%x", e); return -1; }
    return 0;
}

/* Timed: wall seconds of nsteps x (dispatch + barrier) + final glFinish. */
double lp_time_run(int nsteps) {
    struct timespec a, b;
    glFinish_p();
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int s = 0; s < nsteps; s++) one_step(NULL, 0);
    glFinish_p();
    clock_gettime(CLOCK_MONOTONIC, &b);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}
