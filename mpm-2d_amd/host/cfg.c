/*
 * cfg.c -- the configuration-file reader declared in cfg.h: a small
 * tokenizer and recursive reader for the libconfuse subset the reference's
 * .cfg files use (examples/silo/simulation.cfg, tests/isolin.cfg and the like).
 */
#define _GNU_SOURCE
#include "cfg.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct lexer {
    const char *s;
    size_t pos, len;
    const char *path;
    int line;
    int open_comment;   /* a block comment ran to the end of the text */
} lexer;

enum tok { T_END, T_WORD, T_STRING, T_LBRACE, T_RBRACE, T_EQ, T_COMMA, T_LPAREN, T_RPAREN,
           T_ERR };

static void skip_space(lexer *lx)
{
    while (lx->pos < lx->len) {
        char c = lx->s[lx->pos];
        if (c == '\n') {
            lx->line++;
            lx->pos++;
        } else if (isspace((unsigned char)c)) {
            lx->pos++;
        } else if (c == '#' || (c == '/' && lx->pos + 1 < lx->len && lx->s[lx->pos + 1] == '/')) {
            while (lx->pos < lx->len && lx->s[lx->pos] != '\n') lx->pos++;
        } else if (c == '/' && lx->pos + 1 < lx->len && lx->s[lx->pos + 1] == '*') {
            lx->pos += 2;
            while (lx->pos + 1 < lx->len && !(lx->s[lx->pos] == '*' && lx->s[lx->pos + 1] == '/')) {
                if (lx->s[lx->pos] == '\n') lx->line++;
                lx->pos++;
            }
            if (lx->pos + 1 >= lx->len) {
                lx->open_comment = 1;
                lx->pos = lx->len;
                return;
            }
            lx->pos += 2;
        } else {
            break;
        }
    }
}

/* ${VAR} or ${VAR:-default} -> malloc'd text */
static char *expand_env(const char *w, size_t n)
{
    if (n >= 3 && w[0] == '$' && w[1] == '{' && w[n - 1] == '}') {
        char *body = strndup(w + 2, n - 3);
        char *dflt = strstr(body, ":-");
        if (dflt) {
            *dflt = 0;
            dflt += 2;
        }
        const char *v = getenv(body);
        char *out = strdup(v && *v ? v : (dflt ? dflt : ""));
        free(body);
        return out;
    }
    return strndup(w, n);
}

/* Next token; *text gets a malloc'd copy for words and strings. */
static enum tok next(lexer *lx, char **text)
{
    *text = NULL;
    skip_space(lx);
    if (lx->open_comment) return T_ERR;
    if (lx->pos >= lx->len) return T_END;
    char c = lx->s[lx->pos];
    switch (c) {
    case '{': lx->pos++; return T_LBRACE;
    case '}': lx->pos++; return T_RBRACE;
    case '=': lx->pos++; return T_EQ;
    case ',': lx->pos++; return T_COMMA;
    case '(': lx->pos++; return T_LPAREN;
    case ')': lx->pos++; return T_RPAREN;
    default: break;
    }
    if (c == '"' || c == '\'') {
        size_t start = ++lx->pos;
        char *out = malloc(lx->len - start + 1);
        size_t k = 0;
        while (lx->pos < lx->len && lx->s[lx->pos] != c) {
            if (lx->s[lx->pos] == '\\' && lx->pos + 1 < lx->len) lx->pos++;
            if (lx->s[lx->pos] == '\n') lx->line++;
            out[k++] = lx->s[lx->pos++];
        }
        if (lx->pos >= lx->len) {
            free(out);
            return T_ERR;
        }
        lx->pos++;
        out[k] = 0;
        *text = out;
        return T_STRING;
    }
    size_t start = lx->pos;
    if (c == '$' && lx->pos + 1 < lx->len && lx->s[lx->pos + 1] == '{') {
        while (lx->pos < lx->len && lx->s[lx->pos] != '}') lx->pos++;
        if (lx->pos < lx->len) lx->pos++;
    } else {
        while (lx->pos < lx->len) {
            char d = lx->s[lx->pos];
            if (isspace((unsigned char)d) || strchr("{}=,()#\"'", d)) break;
            lx->pos++;
        }
    }
    if (lx->pos == start) return T_ERR;
    *text = expand_env(lx->s + start, lx->pos - start);
    return T_WORD;
}

static void add_entry(cfg_file *cfg, const char *section, const char *key, char **values,
                      size_t count, int is_list)
{
    char full[512];
    snprintf(full, sizeof full, "%s|%s", section, key);
    for (size_t i = 0; i < cfg->count; i++) {
        if (!strcmp(cfg->entries[i].key, full)) {   /* a later setting wins */
            for (size_t k = 0; k < cfg->entries[i].count; k++) free(cfg->entries[i].values[k]);
            free(cfg->entries[i].values);
            cfg->entries[i].values = values;
            cfg->entries[i].count = count;
            cfg->entries[i].is_list = is_list;
            return;
        }
    }
    if (cfg->count == cfg->cap) {
        cfg->cap = cfg->cap ? 2 * cfg->cap : 32;
        cfg->entries = realloc(cfg->entries, cfg->cap * sizeof(cfg_entry));
    }
    cfg_entry *e = &cfg->entries[cfg->count++];
    e->key = strdup(full);
    e->values = values;
    e->count = count;
    e->is_list = is_list;
}

static int fail(cfg_file *cfg, const lexer *lx, const char *what)
{
    snprintf(cfg->error, sizeof cfg->error, "%s:%d: %s", lx->path, lx->line, what);
    return -1;
}

static int read_text(cfg_file *cfg, const char *path, const char *section, int depth);

/* Reads `key = value` and `name { ... }` items until `}` (or end at top). */
static int read_items(cfg_file *cfg, lexer *lx, const char *section, int depth)
{
    for (;;) {
        char *word;
        enum tok t = next(lx, &word);
        if (t == T_END) {
            return section[0] ? fail(cfg, lx, "unterminated section") : 0;
        }
        if (t == T_RBRACE) {
            return section[0] ? 0 : fail(cfg, lx, "unexpected '}'");
        }
        if (t != T_WORD) {
            free(word);
            return fail(cfg, lx, lx->open_comment ? "unterminated comment"
                                                  : "expected an option name");
        }
        char *val;
        enum tok t2 = next(lx, &val);
        if (t2 == T_LPAREN && !strcmp(word, "include")) {
            char *file;
            enum tok t3 = next(lx, &file);
            char *dummy;
            enum tok t4 = next(lx, &dummy);
            free(word);
            free(dummy);
            if ((t3 != T_STRING && t3 != T_WORD) || t4 != T_RPAREN) {
                free(file);
                return fail(cfg, lx, "bad include()");
            }
            int rc = read_text(cfg, file, section, depth + 1);
            free(file);
            if (rc) return rc;
        } else if (t2 == T_LBRACE) {
            int rc = read_items(cfg, lx, word, depth);
            free(word);
            if (rc) return rc;
        } else if (t2 == T_EQ) {
            char *v;
            enum tok t3 = next(lx, &v);
            if (t3 == T_LBRACE) {
                char **vals = NULL;
                size_t n = 0;
                for (;;) {
                    char *item;
                    enum tok ti = next(lx, &item);
                    if (ti == T_RBRACE) break;
                    if (ti == T_COMMA) continue;
                    if (ti != T_WORD && ti != T_STRING) {
                        free(item);
                        for (size_t k = 0; k < n; k++) free(vals[k]);
                        free(vals);
                        free(word);
                        return fail(cfg, lx, "bad list");
                    }
                    vals = realloc(vals, (n + 1) * sizeof(char *));
                    vals[n++] = item;
                }
                add_entry(cfg, section, word, vals, n, 1);
            } else if (t3 == T_WORD || t3 == T_STRING) {
                char **vals = malloc(sizeof(char *));
                vals[0] = v;
                add_entry(cfg, section, word, vals, 1, 0);
            } else {
                free(v);
                free(word);
                return fail(cfg, lx, "expected a value");
            }
            free(word);
        } else {
            free(val);
            free(word);
            return fail(cfg, lx, "expected '=' or '{'");
        }
        free(val);
    }
}

static int read_text(cfg_file *cfg, const char *path, const char *section, int depth)
{
    if (depth > 8) {
        snprintf(cfg->error, sizeof cfg->error, "%s: include nesting too deep", path);
        return -1;
    }
    FILE *fp = fopen(path, "r");
    if (!fp) {
        snprintf(cfg->error, sizeof cfg->error, "cannot open '%s'", path);
        return -1;
    }
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    char *buf = malloc((size_t)n + 1);
    size_t got = fread(buf, 1, (size_t)n, fp);
    fclose(fp);
    buf[got] = 0;
    lexer lx = {buf, 0, got, path, 1, 0};
    int rc = read_items(cfg, &lx, section, depth);
    free(buf);
    return rc;
}

int cfg_read(cfg_file *cfg, const char *path)
{
    memset(cfg, 0, sizeof *cfg);
    const int rc = read_text(cfg, path, "", 0);
    if (rc) {
        /* a file that does not parse leaves no entries, only the message */
        char error[sizeof cfg->error];
        memcpy(error, cfg->error, sizeof error);
        cfg_release(cfg);
        memcpy(cfg->error, error, sizeof error);
    }
    return rc;
}

void cfg_release(cfg_file *cfg)
{
    for (size_t i = 0; i < cfg->count; i++) {
        for (size_t k = 0; k < cfg->entries[i].count; k++) free(cfg->entries[i].values[k]);
        free(cfg->entries[i].values);
        free(cfg->entries[i].key);
    }
    free(cfg->entries);
    memset(cfg, 0, sizeof *cfg);
}

static const cfg_entry *find(const cfg_file *cfg, const char *key)
{
    for (size_t i = 0; i < cfg->count; i++) {
        if (!strcmp(cfg->entries[i].key, key)) return &cfg->entries[i];
    }
    return NULL;
}

const char *cfg_str(const cfg_file *cfg, const char *key, const char *dflt)
{
    const cfg_entry *e = find(cfg, key);
    return e && e->count ? e->values[0] : dflt;
}

double cfg_float(const cfg_file *cfg, const char *key, double dflt)
{
    const char *s = cfg_str(cfg, key, NULL);
    return s ? strtod(s, NULL) : dflt;
}

long cfg_int(const cfg_file *cfg, const char *key, long dflt)
{
    const char *s = cfg_str(cfg, key, NULL);
    return s ? strtol(s, NULL, 0) : dflt;
}

size_t cfg_size(const cfg_file *cfg, const char *key)
{
    const cfg_entry *e = find(cfg, key);
    return e ? e->count : 0;
}

double cfg_nfloat(const cfg_file *cfg, const char *key, size_t i)
{
    const cfg_entry *e = find(cfg, key);
    return e && i < e->count ? strtod(e->values[i], NULL) : 0.0;
}

long cfg_nint(const cfg_file *cfg, const char *key, size_t i)
{
    const cfg_entry *e = find(cfg, key);
    return e && i < e->count ? strtol(e->values[i], NULL, 0) : 0;
}
