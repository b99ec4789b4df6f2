/*
 * cih_stream.h -- the canvas command stream ("wire format").
 *
 * A command stream is a flat little-endian sequence of 32-bit words that
 * records calls on the public canvas API, one record per call, in call order.
 * It is what the serving loop receives per request, what the batch entry point
 * takes per canvas, and what the seeded scene generator emits, so that the
 * same stream can be replayed into any implementation of the canvas API.
 *
 * Layout:
 *   header  : CIH_STREAM_MAGIC, CIH_STREAM_VERSION, width, height
 *   records : opcode word, then the opcode's arguments (floats are stored as
 *             their IEEE-754 bit patterns, ints as two's complement).
 *
 * Each opcode mirrors one public method of canvas_ity::canvas
 * (canvas_ity.hpp:194-1148); the argument order is the method's.  Images are
 * defined once with CIH_OP_DEFINE_IMAGE (id, width, height, then width*height
 * RGBA8 texels packed four bytes per word) and referenced by id from
 * CIH_OP_SET_PATTERN, CIH_OP_DRAW_IMAGE and CIH_OP_PUT_IMAGE_DATA.  Font
 * and text records carry their byte count first and their bytes last, packed
 * four to a word in memory order and zero-padded to a whole word: a font's
 * TrueType file (count 0 with set_font's null font, which only sets the size)
 * and a text's UTF-8 bytes without the terminating NUL (the text ends at its
 * first NUL byte, as a C string does).
 * CIH_OP_SET_FILL_RULE is an extension (the reference only fills non-zero).
 */
#ifndef CIH_STREAM_H
#define CIH_STREAM_H

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CIH_STREAM_MAGIC 0x53484943u /* "CIHS" */
#define CIH_STREAM_VERSION 1u
#define CIH_STREAM_MAX_IMAGES 4096
#define CIH_STREAM_MAX_BLOB (64 << 20) /* font and text records, bytes */

enum cih_opcode {
    CIH_OP_SCALE = 1,              /* f x, f y                     hpp:216  */
    CIH_OP_ROTATE = 2,             /* f angle                      hpp:229  */
    CIH_OP_TRANSLATE = 3,          /* f x, f y                     hpp:241  */
    CIH_OP_TRANSFORM = 4,          /* f a..f                       hpp:262  */
    CIH_OP_SET_TRANSFORM = 5,      /* f a..f                       hpp:290  */
    CIH_OP_SET_GLOBAL_ALPHA = 6,   /* f alpha                      hpp:309  */
    CIH_OP_SET_COMPOSITE_OP = 7,   /* i op (field assignment)      hpp:330  */
    CIH_OP_SET_SHADOW_COLOR = 8,   /* f r,g,b,a                    hpp:350  */
    CIH_OP_SET_SHADOW_BLUR = 9,    /* f level                      hpp:379  */
    CIH_OP_SET_SHADOW_OFFSET = 10, /* f x, f y (fields)            hpp:361  */
    CIH_OP_SET_LINE_WIDTH = 11,    /* f width                      hpp:392  */
    CIH_OP_SET_LINE_CAP = 12,      /* i cap (field assignment)     hpp:404  */
    CIH_OP_SET_LINE_JOIN = 13,     /* i join (field assignment)    hpp:415  */
    CIH_OP_SET_MITER_LIMIT = 14,   /* f limit                      hpp:429  */
    CIH_OP_SET_LINE_DASH = 15,     /* i count (-1 = null), f[count] hpp:460 */
    CIH_OP_SET_LINE_DASH_OFFSET = 16, /* f offset (field)          hpp:440  */
    CIH_OP_SET_COLOR = 17,         /* i type, f r,g,b,a            hpp:478  */
    CIH_OP_SET_LINEAR_GRADIENT = 18, /* i type, f sx,sy,ex,ey      hpp:500  */
    CIH_OP_SET_RADIAL_GRADIENT = 19, /* i type, f sx,sy,sr,ex,ey,er hpp:526 */
    CIH_OP_ADD_COLOR_STOP = 20,    /* i type, f offset,r,g,b,a     hpp:559  */
    CIH_OP_SET_PATTERN = 21,       /* i type, i image, i repetition hpp:592 */
    CIH_OP_BEGIN_PATH = 22,        /*                              hpp:607  */
    CIH_OP_MOVE_TO = 23,           /* f x, f y                     hpp:617  */
    CIH_OP_CLOSE_PATH = 24,        /*                              hpp:629  */
    CIH_OP_LINE_TO = 25,           /* f x, f y                     hpp:642  */
    CIH_OP_QUADRATIC_CURVE_TO = 26, /* f cx,cy,x,y                 hpp:662  */
    CIH_OP_BEZIER_CURVE_TO = 27,   /* f c1x,c1y,c2x,c2y,x,y        hpp:687  */
    CIH_OP_ARC_TO = 28,            /* f vx,vy,x,y,radius           hpp:721  */
    CIH_OP_ARC = 29,               /* f x,y,r,start,end, i ccw     hpp:749  */
    CIH_OP_RECTANGLE = 30,         /* f x,y,w,h                    hpp:771  */
    CIH_OP_FILL = 31,              /*                              hpp:793  */
    CIH_OP_STROKE = 32,            /*                              hpp:813  */
    CIH_OP_CLIP = 33,              /*                              hpp:828  */
    CIH_OP_CLEAR_RECTANGLE = 34,   /* f x,y,w,h                    hpp:862  */
    CIH_OP_FILL_RECTANGLE = 35,    /* f x,y,w,h                    hpp:882  */
    CIH_OP_STROKE_RECTANGLE = 36,  /* f x,y,w,h                    hpp:904  */
    CIH_OP_DRAW_IMAGE = 37,        /* i image, f x,y,w,h           hpp:1057 */
    CIH_OP_SAVE = 38,              /*                              hpp:1141 */
    CIH_OP_RESTORE = 39,           /*                              hpp:1148 */
    CIH_OP_DEFINE_IMAGE = 40,      /* i id, i w, i h, texels       --       */
    CIH_OP_PUT_IMAGE_DATA = 41,    /* i image, i x, i y            hpp:1123 */
    CIH_OP_SET_FILL_RULE = 42,     /* i rule (0 nonzero, 1 evenodd) ext.    */
    CIH_OP_SET_FONT = 43,          /* i bytes, f size, font bytes  hpp:955  */
    CIH_OP_FILL_TEXT = 44,         /* i bytes, f x,y,max, text     hpp:980  */
    CIH_OP_STROKE_TEXT = 45,       /* i bytes, f x,y,max, text     hpp:1006 */
    CIH_OP_SET_TEXT_ALIGN = 46,    /* i align (field assignment)   hpp:924  */
    CIH_OP_SET_TEXT_BASELINE = 47, /* i baseline (field)           hpp:937  */
    CIH_OP_COUNT_
};

/* Number of fixed argument words per opcode; -1 marks variable length. */
static inline int cih_op_arity(uint32_t op)
{
    static const signed char arity[CIH_OP_COUNT_] = {
        0,  2, 1, 2, 6, 6, 1, 1, 4, 1, 2,   /*  0..10 */
        1,  1, 1, 1, -1, 1, 5, 5, 7, 6, 3,  /* 11..21 */
        0,  2, 0, 2, 4, 6, 5, 6, 4, 0, 0,   /* 22..32 */
        0,  4, 4, 4, 5, 0, 0, -1, 3, 1,     /* 33..42 */
        -1, -1, -1, 1, 1 };                 /* 43..47 */
    return op > 0 && op < CIH_OP_COUNT_ ? arity[op] : -2;
}

static inline uint32_t cih_f2w(float value)
{
    uint32_t word;
    memcpy(&word, &value, 4);
    return word;
}

static inline float cih_w2f(uint32_t word)
{
    float value;
    memcpy(&value, &word, 4);
    return value;
}

/* A decoded record: opcode plus a pointer to its argument words. */
typedef struct cih_record {
    uint32_t op;
    const uint32_t *args;
    uint32_t nargs;
} cih_record;

typedef struct cih_reader {
    const uint32_t *words;
    size_t count;
    size_t at;
    int width, height;
} cih_reader;

/* Returns 0 on success, -1 for a malformed header. */
static inline int cih_reader_init(cih_reader *reader, const void *stream,
                                  size_t bytes)
{
    const uint32_t *words = (const uint32_t *)stream;
    if (!stream || bytes < 16 || (bytes & 3) || words[0] != CIH_STREAM_MAGIC ||
        words[1] != CIH_STREAM_VERSION)
        return -1;
    reader->words = words;
    reader->count = bytes / 4;
    reader->at = 4;
    reader->width = (int)words[2];
    reader->height = (int)words[3];
    if (reader->width < 1 || reader->width > 32768 || reader->height < 1 ||
        reader->height > 32768)
        return -1;
    return 0;
}

/* Returns 1 with a record, 0 at the end, -1 on a malformed record. */
static inline int cih_reader_next(cih_reader *reader, cih_record *record)
{
    if (reader->at >= reader->count)
        return 0;
    uint32_t op = reader->words[reader->at];
    int arity = cih_op_arity(op);
    size_t left = reader->count - reader->at - 1;
    const uint32_t *args = reader->words + reader->at + 1;
    size_t nargs;
    if (arity == -2)
        return -1;
    if (arity >= 0)
        nargs = (size_t)arity;
    else if (op == CIH_OP_SET_LINE_DASH) {
        if (left < 1)
            return -1;
        int32_t n = (int32_t)args[0];
        if (n < -1)
            return -1;
        nargs = 1 + (size_t)(n > 0 ? n : 0);
    } else if (op == CIH_OP_SET_FONT || op == CIH_OP_FILL_TEXT ||
               op == CIH_OP_STROKE_TEXT) {
        /* a byte count, then fixed words, then the bytes packed four per word */
        size_t fixed = op == CIH_OP_SET_FONT ? 2 : 4;
        if (left < fixed)
            return -1;
        int32_t n = (int32_t)args[0];
        if (n < 0 || n > CIH_STREAM_MAX_BLOB)
            return -1;
        nargs = fixed + ((size_t)n + 3) / 4;
    } else { /* CIH_OP_DEFINE_IMAGE */
        if (left < 3)
            return -1;
        int32_t w = (int32_t)args[1], h = (int32_t)args[2];
        if (w <= 0 || h <= 0 || w > 32768 || h > 32768 ||
            (int32_t)args[0] < 0 || (int32_t)args[0] >= CIH_STREAM_MAX_IMAGES)
            return -1;
        nargs = 3 + (size_t)w * (size_t)h;
    }
    if (nargs > left)
        return -1;
    record->op = op;
    record->args = args;
    record->nargs = (uint32_t)nargs;
    reader->at += 1 + nargs;
    return 1;
}

#ifdef __cplusplus
}
#endif

#endif /* CIH_STREAM_H */
