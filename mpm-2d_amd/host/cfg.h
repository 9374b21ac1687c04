/*
 * cfg.h -- reader for the configuration files of the reference driver
 * (src/main.c:169-251 declares them with libconfuse, which this image
 * lacks).  It accepts the subset of libconfuse's syntax those files use:
 * sections `name { ... }`, options `key = value`, values that are numbers,
 * quoted strings, bare words or `${VAR:-default}` environment references,
 * lists `{a, b, ...}`, `#`, `//` and C comments, and `include("file")`.
 * Options are kept as text and looked up by "section|key".
 */
#ifndef MPM_CFG_H
#define MPM_CFG_H

#include <stddef.h>

typedef struct cfg_entry {
    char *key;          /* "section|option" */
    char **values;      /* one string per value (lists: per element) */
    size_t count;
    int is_list;
} cfg_entry;

typedef struct cfg_file {
    cfg_entry *entries;
    size_t count, cap;
    char error[256];
} cfg_file;

/* 0 on success; on failure cfg->error holds the message. */
int cfg_read(cfg_file *cfg, const char *path);
void cfg_release(cfg_file *cfg);

/* Lookups with the reference's defaults. */
const char *cfg_str(const cfg_file *cfg, const char *key, const char *dflt);
double cfg_float(const cfg_file *cfg, const char *key, double dflt);
long cfg_int(const cfg_file *cfg, const char *key, long dflt);
/* List option: number of elements (0 if absent) and element i. */
size_t cfg_size(const cfg_file *cfg, const char *key);
double cfg_nfloat(const cfg_file *cfg, const char *key, size_t i);
long cfg_nint(const cfg_file *cfg, const char *key, size_t i);

#endif /* MPM_CFG_H */
