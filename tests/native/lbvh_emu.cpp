/*
 * Sequential CPU emulation of the device linear-BVH build
 * (luxcore_amd/csrc/lrh_gpubuild.hip): the same per-element steps
 * (luxcore_amd/csrc/lrh_lbvh.h) in the order the kernels run them, the same
 * stable LSD radix sort (per-tile digit counts, an exclusive scan of the
 * [digit][tile] table, a scatter in index order), the same error word.  The
 * device build must produce these records byte for byte.
 *
 *   lbvh_emu <in.bin> <outdir> <maxLeaf> <layout> [mode] [method] [travCost]
 * in.bin: u32 nMeshes, then per mesh u32 nVerts, u32 nTris, float[3 nVerts],
 * u32[3 nTris].  mode: "normal" (default); "unsorted": the sort is skipped
 * (codes left in triangle order), as a sort that returned garbage would
 * leave them; "reversed": the sorted codes reversed.  method: "lbvh"
 * (default, the radix tree) or "ploc" (clustering rounds, each round's
 * steps over all clusters before the next step, as the kernels run them).
 * travCost (default 1): the node cost of the leaf decisions.
 * Writes nodes.bin, tris.bin and info.txt ("err <word> nodes <count> depth
 * <d> rounds <r>"); with codes out of order the run stops after the step
 * that reports them, as the device build does, and still terminates: every
 * search is bounded.
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

// the clustering's tree in the radix tree's form, as the device build's
// rounds and conversion make it; returns the rounds, 0 when a round's counts
// are inconsistent
uint32_t plocTree(const std::vector<GBox> &boxes, uint32_t maxLeaf, float ci, std::vector<uint32_t> &vals,
		std::vector<GBox> &leafBox, std::vector<uint32_t> &left2, std::vector<uint32_t> &right2,
		std::vector<uint32_t> &first2, std::vector<uint32_t> &last2, std::vector<GBox> &nodeBox2,
		std::vector<uint32_t> &leafOK2) {
	const uint32_t n = (uint32_t)vals.size(), ni = n - 1;
	std::vector<uint32_t> cid(n), cidB(n), nn(n), lead(n + 1), keep(n + 1);
	std::vector<GBox> cbox(n), cboxB(n), nodeBox(ni);
	std::vector<uint32_t> left(ni), right(ni);
	for (uint32_t i = 0; i < n; ++i) {
		leafBox[i] = boxes[vals[i]];
		cid[i] = kLeafBit | i;
		cbox[i] = leafBox[i];
	}
	std::vector<uint32_t> roundStart{0};
	uint32_t nc = n, nodeBase = 0;
	bool forced = false;
	while (nc > 1) {
		if (roundStart.size() > 4096) return 0;
		if (!forced)
			for (uint32_t i = 0; i < nc; ++i) nn[i] = plocNearest(cbox.data(), nc, i, kPlocRadius);
		for (uint32_t i = 0; i < nc; ++i) {
			const uint32_t j = plocPartner(nn.data(), nc, i, forced);
			lead[i] = (j != kNone && i < j) ? 1u : 0u;
			keep[i] = (j != kNone && i > j) ? 0u : 1u;
		}
		lead[nc] = keep[nc] = 0;
		std::vector<uint32_t> leadOff(lead.begin(), lead.begin() + nc + 1), keepOff(keep.begin(), keep.begin() + nc + 1);
		exclusiveScan(leadOff);
		exclusiveScan(keepOff);
		for (uint32_t i = 0; i < nc; ++i) {
			const uint32_t j = plocPartner(nn.data(), nc, i, forced);
			if (j != kNone && i > j) continue;
			const uint32_t o = keepOff[i];
			if (j != kNone) {
				const uint32_t node = nodeBase + leadOff[i];
				const GBox u = unionBox(cbox[i], cbox[j]);
				left[node] = cid[i];
				right[node] = cid[j];
				nodeBox[node] = u;
				cidB[o] = node;
				cboxB[o] = u;
			} else {
				cidB[o] = cid[i];
				cboxB[o] = cbox[i];
			}
		}
		const uint32_t merged = leadOff[nc], kept = keepOff[nc];
		if (merged == 0 || merged > nc / 2 || kept != nc - merged || nodeBase + merged > ni) return 0;
		nodeBase += merged;
		roundStart.push_back(nodeBase);
		forced = !forced && (uint64_t)merged * 64u < nc;
		nc = kept;
		cid.swap(cidB);
		cbox.swap(cboxB);
	}
	if (nodeBase != ni) return 0;
	std::vector<uint32_t> count(ni), first(ni), leafFirst(n), leafOK(ni);
	std::vector<float> cost(ni);
	for (uint32_t x = 0; x < ni; ++x) {
		const uint32_t l = left[x], r = right[x];
		count[x] = plocSize(l, count.data()) + plocSize(r, count.data());
		const float cl = (l & kLeafBit) ? areaOf(leafBox[l & ~kLeafBit]) : cost[l];
		const float cr = (r & kLeafBit) ? areaOf(leafBox[r & ~kLeafBit]) : cost[r];
		cost[x] = nodeCost(areaOf(nodeBox[x]), cl, cr, count[x], maxLeaf, ci, leafOK[x]);
	}
	first[ni - 1] = 0;
	for (uint32_t x = ni; x-- > 0;) {
		const uint32_t f = first[x], l = left[x], r = right[x], fr = f + plocSize(l, count.data());
		if (l & kLeafBit) leafFirst[l & ~kLeafBit] = f; else first[l] = f;
		if (r & kLeafBit) leafFirst[r & ~kLeafBit] = fr; else first[r] = fr;
	}
	std::vector<uint32_t> vals2(n);
	std::vector<GBox> leafBox2(n);
	for (uint32_t i = 0; i < n; ++i) {
		vals2[leafFirst[i]] = vals[i];
		leafBox2[leafFirst[i]] = leafBox[i];
	}
	auto remap = [&](uint32_t c) { return (c & kLeafBit) ? (kLeafBit | leafFirst[c & ~kLeafBit]) : n - 2 - c; };
	for (uint32_t i = 0; i < ni; ++i) {
		const uint32_t y = n - 2 - i;
		left2[y] = remap(left[i]);
		right2[y] = remap(right[i]);
		first2[y] = first[i];
		last2[y] = first[i] + count[i] - 1;
		nodeBox2[y] = nodeBox[i];
		leafOK2[y] = leafOK[i];
	}
	vals.swap(vals2);
	leafBox.swap(leafBox2);
	return (uint32_t)roundStart.size() - 1;
}

}  // namespace

int main(int argc, char **argv) {
	if (argc < 5) {
		std::cerr << "usage: lbvh_emu in.bin outdir maxLeaf layout [normal|unsorted|reversed] [lbvh|ploc] [travCost]\n";
		return 2;
	}
	const std::string outDir = argv[2];
	const uint32_t maxLeaf = (uint32_t)std::stoul(argv[3]);
	const uint32_t layout = (uint32_t)std::stoul(argv[4]);
	const std::string mode = argc > 5 ? argv[5] : "normal";
	const bool ploc = argc > 6 && std::string(argv[6]) == "ploc";
	const float ci = argc > 7 ? std::stof(argv[7]) : 1.f;
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
	uint32_t err = 0, depth = 0, rounds = 0;
	const size_t nodeSz = layout == 0 ? sizeof(lrh_node4q) : layout == 2 ? sizeof(lrh_node8q) : sizeof(lrh_node4f);
	const int width = layout == 2 ? 8 : 4;
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
		uint32_t ref[kMaxWide];
		for (int k = 0; k < kMaxWide; ++k) ref[k] = k ? LRH_CHILD_EMPTY : leafRef(0, n);
		nodes.assign(nodeSz, 0);
		if (layout == 0)
			lrhdev::quantizeNode(&b, 1, ref, reinterpret_cast<lrh_node4q *>(nodes.data()));
		else if (layout == 2)
			lrhdev::quantizeNode8(&b, 1, ref, reinterpret_cast<lrh_node8q *>(nodes.data()));
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
		std::vector<GBox> nodeBox(ni);
		std::vector<uint32_t> leafOK(ni, 0);
		std::vector<float> cost(ni, 0.f);
		if (ploc && !err) {
			rounds = plocTree(boxes, maxLeaf, ci, vals, leafBox, left, right, first, last, nodeBox, leafOK);
			if (rounds == 0) err |= kErrExpand;
		}
		for (uint32_t i = 0; i < ni && !ploc; ++i) {
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
			if (!ploc)
				for (uint32_t j = 0; j < n; ++j) leafBox[j] = boxes[vals[j]];
			for (uint32_t j = 0; j < n && !ploc; ++j) {
				uint32_t p = lparent[j];
				for (int steps = 0; p < n - 1; ++steps) {
					if (steps >= 128) {
						err |= kErrClimb;
						break;
					}
					if (flags[p]++ == 0u) break;
					const uint32_t l = left[p], r = right[p];
					const GBox &a = (l & kLeafBit) ? leafBox[l & ~kLeafBit] : nodeBox[l];
					const GBox &b = (r & kLeafBit) ? leafBox[r & ~kLeafBit] : nodeBox[r];
					const GBox u = unionBox(a, b);
					const float cl = (l & kLeafBit) ? areaOf(a) : cost[l];
					const float cr = (r & kLeafBit) ? areaOf(b) : cost[r];
					cost[p] = nodeCost(areaOf(u), cl, cr, last[p] - first[p] + 1u, maxLeaf, ci, leafOK[p]);
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
				std::vector<uint32_t> kids((size_t)kMaxWide * nf), cnt(nf + 1, 0);
				for (uint32_t e = 0; e < nf; ++e) {
					if (fA[e] >= ni) {
						err |= kErrExpand;
						continue;
					}
					cnt[e] = expandNode(fA[e], left.data(), right.data(), leafOK.data(), nodeBox.data(),
							width, &kids[(size_t)kMaxWide * e]);
				}
				exclusiveScan(cnt);
				fB.assign(cnt[nf], 0);
				for (uint32_t e = 0; e < nf; ++e)
					emitNode(&kids[(size_t)kMaxWide * e], cnt[e], first.data(), last.data(), leafOK.data(), leafBox.data(),
							nodeBox.data(), levelBase + nf, layout, nodes.data(), levelBase + e, fB.data());
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
	info << "err " << err << " nodes " << nodeCount << " depth " << (depth ? depth - 1 : 0) << " rounds " << rounds
		 << "\n";
	std::cout << "err " << err << " nodes " << nodeCount << " depth " << (depth ? depth - 1 : 0) << " rounds "
			  << rounds << std::endl;
	return 0;
}
