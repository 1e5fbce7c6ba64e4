/*
 * or_tree.h -- the canonical reduction order of the oracle (test
 * infrastructure): values are accumulated in leaves of OR_LEAF consecutive
 * elements from +0.0, and the leaves are summed by a pairwise binary tree,
 * zero padded to a power of two.  tree_push / tree_final keep one partial sum
 * per tree level, so the tree is formed in one pass over the leaves.
 *
 * Any aligned block of 2^k leaves is a complete subtree: summing such blocks
 * separately (in parallel) and pushing the block sums, in order, into a second
 * accumulator gives the same bits as pushing every leaf into one accumulator.
 */
#ifndef OR_TREE_H
#define OR_TREE_H

#include <stdint.h>

#define OR_LEAF 4

typedef struct {
    double stack[64];
    int64_t count;
} tree_acc;

static inline void tree_push(tree_acc *t, double v) {
    int64_t c = t->count;
    int lvl = 0;
    while (c & 1) {
        v = t->stack[lvl] + v;
        c >>= 1;
        ++lvl;
    }
    t->stack[lvl] = v;
    t->count++;
}

static inline double tree_final(const tree_acc *t) {
    double acc = 0.0;
    int have = 0;
    for (int lvl = 0; lvl < 63; ++lvl) {
        if ((t->count >> lvl) & 1) {
            acc = have ? t->stack[lvl] + acc : t->stack[lvl];
            have = 1;
        }
    }
    return have ? acc : 0.0;
}

#endif
