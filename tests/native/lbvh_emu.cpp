/*
 * Sequential CPU emulation of the device linear-BVH build
 * (luxcore_amd/csrc/lrh_gpubuild.hip): the same per-element steps
 * (luxcore_amd/csrc/lrh_lbvh.h) in the order the kernels run them, the same
 * stable LSD radix sort (per-tile digit counts, an exclusive scan of the
 * [digit][tile] table, a scatter in index order), the same error word.  The
 * device build must produce these records byte for byte.
 *
 *   lbvh_emu <in.bin> <outdir> <maxLeaf> <layout> [mode]
 * in.bin: u32 nMeshes, then per mesh u32 nVerts, u32 nTris, float[3 nVerts],
 * u32[3 nTris].  mode: "normal" (default); "unsorted": the sort is skipped
 * (codes left in triangle order), as a sort that returned garbage would
 * leave them; "reversed": the sorted codes reversed.  Writes nodes.bin,
 * tris.bin and info.txt ("err <word> nodes <count> depth <d>"); with
 * codes out of order the run stops after the step that reports them, as the
 * device build does, and still terminates: every search is bounded.
 */
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../luxcore_amd/csrc/lrh_lbvh.h"

using namespace lrhlbvh;

namespace {

template <class T>
void put(const std::string &path, const std::vector<T> &v) {
	std::ofstream f(path, std::ios::binary);
	f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

void exclusiveScan(std::vector<uint32_t> &v) {
	uint32_t run = 0;
	for (uint32_t &x : v) {
		const uint32_t a = x;
		x = run;
		run += a;
	}
}

}  // namespace

int main(int argc, char **argv) {
	if (argc < 5) {
		std::cerr << "usage: lbvh_emu in.bin outdir maxLeaf layout [normal|unsorted|reversed]\n";
		return 2;
	}
	const std::string outDir = argv[2];
	const uint32_t maxLeaf = (uint32_t)std::stoul(argv[3]);
	const uint32_t layout = (uint32_t)std::stoul(argv[4]);
	const std::string mode = argc > 5 ? argv[5] : "normal";
	std::ifstream in(argv[1], std::ios::binary);
	auto rd = [&](void *p, size_t b) {
		in.read(reinterpret_cast<char *>(p), (std::streamsize)b);
		if (!in) throw std::runtime_error("short input");
	};
	uint32_t nMeshes = 0;
	rd(&nMeshes, 4);
	std::vector<float> verts;
	std::vector<uint32_t> tris, tbase(1, 0), vbase(1, 0);
	for (uint32_t m = 0; m < nMeshes; ++m) {
		uint32_t nv = 0, nt = 0;
		rd(&nv, 4);
		rd(&nt, 4);
		const size_t v0 = verts.size(), t0 = tris.size();
		verts.resize(v0 + 3 * (size_t)nv);
		tris.resize(t0 + 3 * (size_t)nt);
		rd(verts.data() + v0, 12 * (size_t)nv);
		rd(tris.data() + t0, 12 * (size_t)nt);
		tbase.push_back(tbase.back() + nt);
		vbase.push_back(vbase.back() + nv);
	}
	const uint32_t n = tbase.back();
	uint32_t err = 0, depth = 0;
	const size_t nodeSz = layout == 0 ? sizeof(lrh_node4q) : sizeof(lrh_node4f);
	std::vector<uint8_t> nodes;
	std::vector<lrh_tri> recs;
	std::vector<uint32_t> vals(n);
	std::vector<GBox> leafBox(n);
	// 1. leaf boxes and centre bounds
	std::vector<GBox> boxes(n);
	uint32_t cb[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
	for (uint32_t i = 0; i < n; ++i) {
		boxes[i] = primBox(verts.data(), tris.data(), tbase.data(), vbase.data(), nMeshes, i);
		for (int k = 0; k < 3; ++k) {
			const uint32_t c = lrhdev::fkey(boxes[i].lo[k] + boxes[i].hi[k]);
			cb[k] = std::min(cb[k], c);
			cb[3 + k] = std::max(cb[3 + k], c);
		}
	}
	uint32_t nodeCount = 0;
	if (n <= maxLeaf) {
		for (uint32_t i = 0; i < n; ++i) {
			vals[i] = i;
			leafBox[i] = boxes[i];
		}
		GBox b;
		for (int k = 0; k < 3; ++k) {
			b.lo[k] = INFINITY;
			b.hi[k] = -INFINITY;
		}
		for (uint32_t i = 0; i < n; ++i)
			for (int k = 0; k < 3; ++k) {
				b.lo[k] = minT(b.lo[k], leafBox[i].lo[k]);
				b.hi[k] = maxT(b.hi[k], leafBox[i].hi[k]);
			}
		const uint32_t ref[4] = {leafRef(0, n), LRH_CHILD_EMPTY, LRH_CHILD_EMPTY, LRH_CHILD_EMPTY};
		nodes.assign(nodeSz, 0);
		if (layout == 0)
			lrhdev::quantizeNode(&b, 1, ref, reinterpret_cast<lrh_node4q *>(nodes.data()));
		else
			lrhdev::floatNode(&b, 1, ref, reinterpret_cast<lrh_node4f *>(nodes.data()));
		nodeCount = 1;
	} else {
		// 2. Morton codes and the radix sort
		std::vector<uint64_t> keys(n), keysB(n);
		std::vector<uint32_t> valsB(n);
		for (uint32_t i = 0; i < n; ++i) {
			keys[i] = mortonKey(boxes[i], cb);
			vals[i] = i;
		}
		if (mode != "unsorted") {
			const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
			std::vector<uint32_t> table((size_t)256 * tiles);
			for (uint32_t shift = 0; shift < 64; shift += 8) {
				std::fill(table.begin(), table.end(), 0u);
				for (uint32_t i = 0; i < n; ++i) ++table[(size_t)digitOf(keys[i], shift) * tiles + i / kSortTile];
				exclusiveScan(table);
				for (uint32_t i = 0; i < n; ++i) {
					uint32_t &slot = table[(size_t)digitOf(keys[i], shift) * tiles + i / kSortTile];
					keysB[slot] = keys[i];
					valsB[slot] = vals[i];
					++slot;
				}
				keys.swap(keysB);
				vals.swap(valsB);
			}
			if (mode == "reversed") {
				std::reverse(keys.begin(), keys.end());
				std::reverse(vals.begin(), vals.end());
			}
		}
		for (uint32_t i = 0; i + 1 < n; ++i)
			if (keys[i] > keys[i + 1]) err |= kErrUnsorted;
		// 3. radix tree (run even on out-of-order codes: it must terminate and report them)
		const uint32_t ni = n - 1;
		std::vector<uint32_t> left(ni), right(ni), first(ni), last(ni), iparent(ni, kNone), lparent(n, kNone),
				flags(ni, 0);
		for (uint32_t i = 0; i < ni; ++i) {
			uint32_t lc, rc, lo, hi;
			if (!radixNode(keys.data(), n, i, lc, rc, lo, hi)) {
				err |= kErrRange;
				continue;
			}
			left[i] = lc;
			right[i] = rc;
			first[i] = lo;
			last[i] = hi;
			if (lc & kLeafBit) lparent[lc & ~kLeafBit] = i; else iparent[lc] = i;
			if (rc & kLeafBit) lparent[rc & ~kLeafBit] = i; else iparent[rc] = i;
		}
		if (!err) {
			// 4. boxes bottom-up, the second arrival at a node computes it
			std::vector<GBox> nodeBox(ni);
			for (uint32_t j = 0; j < n; ++j) leafBox[j] = boxes[vals[j]];
			for (uint32_t j = 0; j < n; ++j) {
				uint32_t p = lparent[j];
				for (int steps = 0; p < n - 1; ++steps) {
					if (steps >= 128) {
						err |= kErrClimb;
						break;
					}
					if (flags[p]++ == 0u) break;
					const uint32_t cl = left[p], cr = right[p];
					const GBox &a = (cl & kLeafBit) ? leafBox[cl & ~kLeafBit] : nodeBox[cl];
					const GBox &b = (cr & kLeafBit) ? leafBox[cr & ~kLeafBit] : nodeBox[cr];
					GBox u;
					for (int k = 0; k < 3; ++k) {
						u.lo[k] = minT(a.lo[k], b.lo[k]);
						u.hi[k] = maxT(a.hi[k], b.hi[k]);
					}
					nodeBox[p] = u;
					p = iparent[p];
				}
			}
			// 5. breadth-first collapse
			nodes.assign(std::max<size_t>(ni, 1) * nodeSz, 0);
			std::vector<uint32_t> fA(1, 0), fB;
			uint32_t levelBase = 0;
			while (!fA.empty() && !err) {
				const uint32_t nf = (uint32_t)fA.size();
				if ((size_t)levelBase + nf > ni || depth > 128) {
					err |= kErrExpand;
					break;
				}
				std::vector<uint32_t> kids(4 * (size_t)nf), cnt(nf + 1, 0);
				for (uint32_t e = 0; e < nf; ++e) {
					if (fA[e] >= ni) {
						err |= kErrExpand;
						continue;
					}
					cnt[e] = expandNode(fA[e], left.data(), right.data(), first.data(), last.data(), nodeBox.data(),
							maxLeaf, &kids[4 * (size_t)e]);
				}
				exclusiveScan(cnt);
				fB.assign(cnt[nf], 0);
				for (uint32_t e = 0; e < nf; ++e)
					emitNode(&kids[4 * (size_t)e], cnt[e], first.data(), last.data(), leafBox.data(), nodeBox.data(),
							maxLeaf, levelBase + nf, layout, nodes.data(), levelBase + e, fB.data());
				levelBase += nf;
				fA.swap(fB);
				++depth;
			}
			nodeCount = levelBase;
			nodes.resize((size_t)nodeCount * nodeSz);
		}
	}
	if (!err) {
		// 6. triangle records in leaf order
		recs.resize(n);
		for (uint32_t j = 0; j < n; ++j)
			recs[j] = triRecord(verts.data(), tris.data(), tbase.data(), vbase.data(), nMeshes, vals[j]);
	}
	put(outDir + "/nodes.bin", nodes);
	put(outDir + "/tris.bin", recs);
	std::ofstream info(outDir + "/info.txt");
	info << "err " << err << " nodes " << nodeCount << " depth " << (depth ? depth - 1 : 0) << "\n";
	std::cout << "err " << err << " nodes " << nodeCount << " depth " << (depth ? depth - 1 : 0) << std::endl;
	return 0;
}
