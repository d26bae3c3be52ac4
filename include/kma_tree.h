/*
 * kma_tree.h - device side of the insert-only binary search tree, HIP for
 * gfx950.  Header-only.
 *
 * Same node layout and semantics as the reference's clTree (clTree.cl:27-74,
 * clTree.h:24-32): kma_tree_add links a caller-owned node under the first
 * NULL child on its key's path with one compare-and-swap per attempt and
 * returns 1, or returns 0 (node not linked) when a node with the same key is
 * already in the tree; kma_tree_get walks from the root and returns the node
 * with the key or NULL.  Keys must not change after a node is added.
 *
 * Visibility: add writes the node's key and NULL children with agent-scope
 * stores and writes back this XCD's L2 once per wavefront before linking,
 * so the node (and whatever else the caller stored into it) is complete for
 * a reader on any XCD; the walks read child links and keys with agent-scope
 * loads.
 */
#ifndef KMA_TREE_H
#define KMA_TREE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef struct kma_tree_node {
	uint32_t key;
	uint32_t pad;
	struct kma_tree_node *left;
	struct kma_tree_node *right;
} kma_tree_node;

typedef struct kma_tree {
	kma_tree_node *root;
} kma_tree;

namespace kma_t {
__device__ __forceinline__ uint64_t ld(kma_tree_node *const *p)
{
	return __hip_atomic_load((const uint64_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ldk(const uint32_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
} // namespace kma_t

__device__ inline uint32_t kma_tree_add(kma_tree *t, kma_tree_node *n)
{
	using namespace kma_t;
	const uint32_t key = n->key;
	__hip_atomic_store((uint64_t *)&n->left, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__hip_atomic_store((uint64_t *)&n->right, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__hip_atomic_store(&n->key, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	kma_tree_node **cursor = &t->root;
	for (;;) {
		uint64_t expect = 0;
		if (__hip_atomic_compare_exchange_strong((uint64_t *)cursor, &expect, (uint64_t)(uintptr_t)n,
							 __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
			return 1;
		kma_tree_node *c = (kma_tree_node *)(uintptr_t)expect;
		uint32_t ck = ldk(&c->key);
		if (ck == key)
			return 0;
		cursor = key < ck ? &c->left : &c->right;
	}
}

__device__ inline kma_tree_node *kma_tree_get(kma_tree *t, uint32_t key)
{
	using namespace kma_t;
	kma_tree_node *c = (kma_tree_node *)(uintptr_t)ld(&t->root);
	while (c) {
		uint32_t ck = ldk(&c->key);
		if (ck == key)
			return c;
		c = (kma_tree_node *)(uintptr_t)ld(key < ck ? &c->left : &c->right);
	}
	return nullptr;
}

#endif /* KMA_TREE_H */
