#ifndef VEC3_H
#define VEC3_H

/*
 * 3-component float vector used by the render data model.
 * Same type and prototypes as the reference's vec3.h (vec3.h:4-33) so that a
 * caller which links its own vec3.c keeps compiling against this include
 * directory.  librend itself does not export these functions.
 */
struct vec3 {
	float  x, y, z;
};

struct vec3  vec3_randuni(unsigned int *seed);

struct vec3  vec3_add(struct vec3 a, struct vec3 b);
struct vec3  vec3_sub(struct vec3 a, struct vec3 b);
struct vec3  vec3_mul(struct vec3 a, struct vec3 b);

struct vec3  vec3_neg(struct vec3 v);
struct vec3  vec3_scale(struct vec3 v, float s);

struct vec3  vec3_cross(struct vec3 a, struct vec3 b);
struct vec3  vec3_unit(struct vec3 v);

float        vec3_dot(struct vec3 a, struct vec3 b);
float        vec3_len(struct vec3 v);

struct vec3  vec3_min(struct vec3 a, struct vec3 b);
struct vec3  vec3_max(struct vec3 a, struct vec3 b);

float        vec3_minc(struct vec3 v);
float        vec3_maxc(struct vec3 v);

float        vec3_getc(struct vec3 v, unsigned char i);

struct vec3  vec3_abs(struct vec3 v);

struct vec3  vec3_spherical(float theta, float phi);

#endif
