/* Embeds the reference's GLSL compute sources into oracle/_ref/libse_llvmpipe.so
 * at build time.  The files are read where they lie (assembler include path set
 * to <reference>/shaders/compute by oracle/Makefile); nothing is copied into the
 * repository.  Each blob is NUL-terminated. */
    .section .rodata
    .p2align 4

#define BLOB(sym, file)          \
    .globl sym;                  \
sym:                             \
    .incbin file;                \
    .byte 0

BLOB(se_glsl_falling_sand, "falling_sand.glsl")
BLOB(se_glsl_operations, "operations.glsl")
BLOB(se_glsl_math, "math.glsl")
BLOB(se_glsl_cell, "cell.glsl")
BLOB(se_glsl_directions, "directions.glsl")
BLOB(se_glsl_material_helpers, "material_helpers.glsl")
BLOB(se_glsl_gen_materials, "gen/materials.glsl")
BLOB(se_glsl_gen_rules, "gen/rules.glsl")
BLOB(se_glsl_gen_falling_sand, "gen/falling_sand.glsl")

    .section .note.GNU-stack,"",@progbits
