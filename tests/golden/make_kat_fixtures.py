"""Writes ``kat_compile.json``: literal known-answer values of the compiled index tensors,
transcribed by hand from the reference's tests (values only):

* ``tests/model/simple_model_test.py:55-189`` and ``:446-488``   -> "simple_model"
* ``tests/layer/sum_layer_test.py:49-79``                        -> "sum_layer"
* ``tests/layer/prod_layer_test.py:38-50``                       -> "prod_layer_bs1", "prod_layer_bs8"
* ``tests/layer/input_layer_test.py:34-39`` and ``:156-167``     -> "input_layer", "input_layer_tied"
* ``tests/model/parameter_tying_test.py:50-150``                 -> "tying_block1"

Each check is ``[expression, expected]``; expressions are evaluated by
``tests/test_compile_kat.py`` against the objects built by ``tests/kat_circuits.py``.
``{"arange": [a, b, step]}`` denotes ``torch.arange(a, b, step)``; ``{"full": v}`` means
every element equals ``v``; ``{"bcast": x}`` means ``x`` broadcast over the compared rows.
"""
import json
import os


def ar(a, b, s=1):
    return {"arange": [a, b, s]}


def full(v):
    return {"full": v}


def bc(x):
    return {"bcast": x}


KATS = {
    "simple_model": [
        ["il.vids[0:32,0]", full(0)], ["il.vids[32:64,0]", full(1)],
        ["il.vids[64:96,0]", full(2)], ["il.vids[96:128,0]", full(3)],
        ["il.s_pids[:64]", ar(0, 256, 4)], ["il.s_pids[64:]", ar(256, 640, 6)],
        ["il.s_pfids[:64]", ar(0, 256, 4)], ["il.s_pfids[64:]", ar(256, 640, 6)],
        ["il.s_mids[0:32]", full(0)], ["il.s_mids[32:64]", full(1)],
        ["il.s_mids[64:96]", full(2)], ["il.s_mids[96:128]", full(3)],
        ["il.source_nids", ar(0, 128)], ["il.num_parameters", 640],
        ["pl0.num_nodes", 128], ["pl0.num_edges", 256],
        ["pl0.partitioned_nids[0]", ar(16, 144, 16)],
        ["pl0.partitioned_cids[0][0:2,:]", [[16, 48], [32, 64]]],
        ["pl0.partitioned_cids[0][2:4,:]", [[16, 48], [32, 64]]],
        ["pl0.partitioned_cids[0][4:6,:]", [[80, 112], [96, 128]]],
        ["pl0.partitioned_cids[0][6:8,:]", [[48, 80], [64, 96]]],
        ["pl0.partitioned_u_cids[0]", [16, 32, 80, 96, 112, 128]],
        ["pl0.partitioned_u_cids[1]", [48, 64]],
        ["pl0.partitioned_parids[0]", [[16, 48], [32, 64], [80, 112], [96, 128], [80, 0], [96, 0]]],
        ["pl0.partitioned_parids[1]", [[16, 48, 112, 0], [32, 64, 128, 0]]],
        ["sl0.partitioned_nids[0]", [176, 192, 208, 224]],
        ["sl0.partitioned_nids[1]", [144, 160]],
        ["sl0.partitioned_cids[0][0,:]", ar(80, 112)], ["sl0.partitioned_cids[0][1,:]", ar(80, 112)],
        ["sl0.partitioned_cids[0][2,:]", ar(112, 144)], ["sl0.partitioned_cids[0][3,:]", ar(112, 144)],
        ["sl0.partitioned_cids[1][0,:]", ar(16, 80)], ["sl0.partitioned_cids[1][1,:]", ar(16, 80)],
        ["sl0.partitioned_pids[0][0,:]", ar(2304, 2816, 16)], ["sl0.partitioned_pids[0][1,:]", ar(2816, 3328, 16)],
        ["sl0.partitioned_pids[0][2,:]", ar(3328, 3840, 16)], ["sl0.partitioned_pids[0][3,:]", ar(3840, 4352, 16)],
        ["sl0.partitioned_pids[1][0,:]", ar(256, 1280, 16)], ["sl0.partitioned_pids[1][1,:]", ar(1280, 2304, 16)],
        ["sl0.partitioned_pfids[0][0,:]", ar(2048, 2560, 16)], ["sl0.partitioned_pfids[0][1,:]", ar(2560, 3072, 16)],
        ["sl0.partitioned_pfids[0][2,:]", ar(3072, 3584, 16)], ["sl0.partitioned_pfids[0][3,:]", ar(3584, 4096, 16)],
        ["sl0.partitioned_pfids[1][0,:]", ar(0, 1024, 16)], ["sl0.partitioned_pfids[1][1,:]", ar(1024, 2048, 16)],
        ["sl0.partitioned_chids[0]", ar(16, 144, 16)],
        ["sl0.partitioned_parids[0][:4]", bc([144, 160])],
        ["sl0.partitioned_parids[0][4:6]", bc([176, 192])],
        ["sl0.partitioned_parids[0][6:8]", bc([208, 224])],
        ["sl0.partitioned_parpids[0][0,:]", [256, 1280]], ["sl0.partitioned_parpids[0][1,:]", [512, 1536]],
        ["sl0.partitioned_parpids[0][2,:]", [768, 1792]], ["sl0.partitioned_parpids[0][3,:]", [1024, 2048]],
        ["sl0.partitioned_parpids[0][4,:]", [2304, 2816]], ["sl0.partitioned_parpids[0][5,:]", [2560, 3072]],
        ["sl0.partitioned_parpids[0][6,:]", [3328, 3840]], ["sl0.partitioned_parpids[0][7,:]", [3584, 4096]],
        ["pl1.partitioned_nids[0]", ar(16, 112, 16)],
        ["pl1.partitioned_cids[0][0,:]", [144, 80, 112, 0]], ["pl1.partitioned_cids[0][1,:]", [160, 96, 128, 0]],
        ["pl1.partitioned_cids[0][2,:]", [176, 16, 48, 0]], ["pl1.partitioned_cids[0][3,:]", [192, 32, 64, 0]],
        ["pl1.partitioned_cids[0][4,:]", [208, 16, 112, 0]], ["pl1.partitioned_cids[0][5,:]", [224, 32, 128, 0]],
        ["pl1.partitioned_u_cids[0]", [48, 64, 80, 96, 144, 160, 176, 192, 208, 224]],
        ["pl1.partitioned_u_cids[1]", [16, 32, 112, 128]],
        ["pl1.partitioned_parids[0][0:2,:]", [[48], [64]]], ["pl1.partitioned_parids[0][2:4,:]", [[16], [32]]],
        ["pl1.partitioned_parids[0][4:6,:]", [[16], [32]]], ["pl1.partitioned_parids[0][6:8,:]", [[48], [64]]],
        ["pl1.partitioned_parids[0][8:10,:]", [[80], [96]]],
        ["pl1.partitioned_parids[1][0:2,:]", [[48, 80], [64, 96]]],
        ["pl1.partitioned_parids[1][2:4,:]", [[16, 80], [32, 96]]],
        ["sl1.block_size", 1], ["sl1.partitioned_nids[0]", [240]],
        ["sl1.partitioned_cids[0][0,:96]", ar(16, 112)], ["sl1.partitioned_cids[0][0,96:]", full(0)],
        ["sl1.partitioned_pids[0][0,:96]", ar(4352, 4448)], ["sl1.partitioned_pids[0][0,96:]", full(0)],
        ["sl1.partitioned_chids[0]", ar(16, 112, 16)], ["sl1.partitioned_parids[0]", full(240)],
        ["sl1.partitioned_parpids[0][:,0]", ar(4352, 4448, 16)],
        ["pc.params[:256]", full(0.0)],
        ["em.par_start_ids[0:16]", ar(256, 272)], ["em.par_start_ids[16:32]", ar(768, 784)],
        ["em.par_start_ids[32:48]", ar(1280, 1296)], ["em.par_start_ids[48:64]", ar(1792, 1808)],
        ["em.par_start_ids[64:80]", ar(2304, 2320)], ["em.par_start_ids[80:96]", ar(2816, 2832)],
        ["em.par_start_ids[96:112]", ar(3328, 3344)], ["em.par_start_ids[112:128]", ar(3840, 3856)],
        ["em.par_start_ids[128:131]", [4352, 4384, 4416]],
        ["em.par_start_ids - em.pflow_start_ids", full(256)],
        ["em.blk_sizes[0:131]", full(32)],
        ["em.blk_intervals[0:128]", full(16)], ["em.blk_intervals[128:131]", full(1)],
        ["em.global_nids[0:16]", ar(0, 16)], ["em.global_nids[16:32]", ar(0, 16)],
        ["em.global_nids[32:48]", ar(16, 32)], ["em.global_nids[48:64]", ar(16, 32)],
        ["em.global_nids[64:80]", ar(32, 48)], ["em.global_nids[80:96]", ar(48, 64)],
        ["em.global_nids[96:112]", ar(64, 80)], ["em.global_nids[112:128]", ar(80, 96)],
        ["em.global_nids[128:131]", full(96)],
        ["em.nchs[0:32]", full(64)], ["em.nchs[32:64]", full(64)], ["em.nchs[64:128]", full(32)],
        ["em.nchs[128:131]", full(96)], ["em.cum_pflows.size(0)", 97],
    ],
    "sum_layer": [
        ["layer.partitioned_nids[0]", ar(16, 112, 16)],
        ["layer.partitioned_cids[0][0:2,0]", full(16)], ["layer.partitioned_cids[0][2:4,0]", full(48)],
        ["layer.partitioned_cids[0][4:6,0]", full(80)], ["layer.partitioned_cids[0][0:2,1]", full(17)],
        ["layer.partitioned_cids[0][2:4,1]", full(49)], ["layer.partitioned_cids[0][4:6,1]", full(81)],
        ["layer.partitioned_pids[0][:,0]", [256, 768, 1280, 1792, 2304, 2816]],
        ["layer.partitioned_pids[0][:,1]", [272, 784, 1296, 1808, 2320, 2832]],
        ["layer.partitioned_pfids[0][:,0]", [0, 512, 1024, 1536, 2048, 2560]],
        ["layer.partitioned_pfids[0][:,1]", [16, 528, 1040, 1552, 2064, 2576]],
        ["layer.partitioned_chids[0]", ar(16, 112, 16)],
        ["layer.partitioned_parids[0][0:2,0]", full(16)], ["layer.partitioned_parids[0][0:2,1]", full(32)],
        ["layer.partitioned_parids[0][2:4,0]", full(48)], ["layer.partitioned_parids[0][2:4,1]", full(64)],
        ["layer.partitioned_parids[0][4:6,0]", full(80)], ["layer.partitioned_parids[0][4:6,1]", full(96)],
        ["layer.partitioned_parpids[0][0,0]", 256], ["layer.partitioned_parpids[0][1,0]", 512],
        ["layer.partitioned_parpids[0][0,1]", 768], ["layer.partitioned_parpids[0][1,1]", 1024],
        ["layer.partitioned_parpids[0][2,0]", 1280], ["layer.partitioned_parpids[0][3,0]", 1536],
        ["layer.partitioned_parpids[0][2,1]", 1792], ["layer.partitioned_parpids[0][3,1]", 2048],
        ["layer.partitioned_parpids[0][4,0]", 2304], ["layer.partitioned_parpids[0][5,0]", 2560],
        ["layer.partitioned_parpids[0][4,1]", 2816], ["layer.partitioned_parpids[0][5,1]", 3072],
    ],
    "prod_layer_bs1": [
        ["layer.partitioned_nids[0]", ar(1, 7, 1)],
        ["layer.partitioned_cids[0][0,:]", [1, 3]], ["layer.partitioned_cids[0][1,:]", [2, 4]],
        ["layer.partitioned_cids[0][2,:]", [5, 7]], ["layer.partitioned_cids[0][3,:]", [6, 8]],
        ["layer.partitioned_cids[0][4,:]", [3, 5]], ["layer.partitioned_cids[0][5,:]", [4, 6]],
    ],
    "prod_layer_bs8": [
        ["layer.partitioned_nids[0]", ar(8, 56, 8)],
        ["layer.partitioned_cids[0][0,:]", [8, 24]], ["layer.partitioned_cids[0][1,:]", [16, 32]],
        ["layer.partitioned_cids[0][2,:]", [40, 56]], ["layer.partitioned_cids[0][3,:]", [48, 64]],
        ["layer.partitioned_cids[0][4,:]", [24, 40]], ["layer.partitioned_cids[0][5,:]", [32, 48]],
    ],
    "input_layer": [
        ["layer.vids[:,0]", [0] * 8 + [1] * 8 + [2] * 8 + [3] * 8],
        ["layer.s_pids", ar(0, 64, 2)], ["layer.s_pfids", ar(0, 64, 2)],
        ["layer.metadata", [2.0, 2.0, 2.0, 2.0]],
        ["layer.s_mids", [0] * 8 + [1] * 8 + [2] * 8 + [3] * 8],
        ["layer.source_nids", ar(0, 32)],
    ],
    "input_layer_tied": [
        ["layer.vids[:,0]", [0] * 8 + [1] * 8 + [2] * 8 + [3] * 8],
        ["layer.s_pids", list(range(0, 48, 2)) + list(range(16, 32, 2))],
        ["layer.s_pfids", ar(0, 64, 2)], ["layer.metadata", [2.0, 2.0, 2.0, 2.0]],
        ["layer.s_mids", [0] * 8 + [1] * 8 + [2] * 8 + [3] * 8],
        ["layer.source_nids", ar(0, 24)],
        ["layer.tied2source_nids[0][0]", 16], ["layer.tied2source_nids[0][1]", 16],
        ["layer.tied2source_nids[0][2]", [16, 48]],
    ],
    "tying_block1": [
        ["il.vids[:,0]", [0, 0, 1, 1, 2, 2, 3, 3]],
        ["il.s_pids", [0, 5, 10, 15, 20, 25, 30, 35]], ["il.s_pfids", [0, 5, 10, 15, 20, 25, 30, 35]],
        ["il.metadata", [5.0, 5.0, 5.0, 5.0]], ["il.s_mids", [0, 0, 1, 1, 2, 2, 3, 3]],
        ["il.source_nids", [0, 1, 2, 3, 4, 5, 6, 7]],
        ["il._output_ind_range[0]", 1], ["il._output_ind_range[1]", 9],
        ["g[0].partitioned_nids[0]", [1, 2, 3, 4, 5, 6]],
        ["g[0].partitioned_cids[0]", [[1, 3], [2, 4], [3, 5], [4, 6], [5, 7], [6, 8]]],
        ["g[0].partitioned_u_cids[0]", [1, 2, 3, 4, 5, 6, 7, 8]],
        ["g[0].partitioned_parids[0]", [[1, 0], [2, 0], [1, 3], [2, 4], [3, 5], [4, 6], [5, 0], [6, 0]]],
        ["g[1].partitioned_nids[0]", ar(9, 15)],
        ["g[1].partitioned_cids[0]", [[1, 2], [1, 2], [3, 4], [3, 4], [5, 6], [5, 6]]],
        ["g[1].partitioned_pids[0]", [[1, 2], [3, 4], [1, 2], [3, 4], [1, 2], [3, 4]]],
        ["g[1].partitioned_pfids[0]", [[0, 1], [2, 3], [0, 1], [2, 3], [4, 5], [6, 7]]],
        ["g[1].partitioned_chids[0]", ar(1, 7)],
        ["g[1].partitioned_parids[0]", [[9, 10], [9, 10], [11, 12], [11, 12], [13, 14], [13, 14]]],
        ["g[1].partitioned_parpids[0]", [[1, 3], [2, 4], [1, 3], [2, 4], [1, 3], [2, 4]]],
        ["g[2].partitioned_nids[0]", ar(1, 9)],
        ["g[2].partitioned_cids[0]", [[9, 5], [10, 6], [11, 1], [12, 2], [11, 7], [12, 8], [13, 3], [14, 4]]],
        ["g[2].partitioned_u_cids[0]", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 13, 14]],
        ["g[2].partitioned_u_cids[1]", [11, 12]],
        ["g[2].partitioned_parids[0][:,0]", [3, 4, 7, 8, 1, 2, 5, 6, 1, 2, 7, 8]],
        ["g[2].partitioned_parids[1]", [[3, 5], [4, 6]]],
        ["g[3].partitioned_nids[0]", ar(15, 19)],
        ["g[3].partitioned_cids[0]", [[1, 2, 3, 4], [1, 2, 3, 4], [5, 6, 7, 8], [5, 6, 7, 8]]],
        ["g[3].partitioned_pids[0]", [[5, 6, 7, 8], [9, 10, 11, 12], [5, 6, 7, 8], [9, 10, 11, 12]]],
        ["g[3].partitioned_pfids[0]", [[8, 9, 10, 11], [12, 13, 14, 15], [8, 9, 10, 11], [12, 13, 14, 15]]],
        ["g[3].partitioned_chids[0]", ar(1, 9)],
        ["g[3].partitioned_parids[0][0:4]", bc([15, 16])], ["g[3].partitioned_parids[0][4:8]", bc([17, 18])],
        ["g[3].partitioned_parpids[0][0:4]", [[5, 9], [6, 10], [7, 11], [8, 12]]],
        ["g[3].partitioned_parpids[0][4:8]", [[5, 9], [6, 10], [7, 11], [8, 12]]],
        ["g[4].partitioned_nids[0]", [1, 2, 3, 4]],
        ["g[4].partitioned_cids[0]", [[15, 7], [16, 8], [17, 1], [18, 2]]],
        ["g[4].partitioned_u_cids[0]", [1, 2, 7, 8, 15, 16, 17, 18]],
        ["g[4].partitioned_parids[0][:,0]", [3, 4, 1, 2, 1, 2, 3, 4]],
        ["g[5].partitioned_nids[0]", ar(19, 21)],
        ["g[5].partitioned_cids[0]", [[1, 2, 3, 4], [1, 2, 3, 4]]],
        ["g[5].partitioned_pids[0]", [[5, 6, 7, 8], [9, 10, 11, 12]]],
        ["g[5].partitioned_pfids[0]", [[8, 9, 10, 11], [12, 13, 14, 15]]],
        ["g[5].partitioned_chids[0]", ar(1, 5)],
        ["g[5].partitioned_parids[0]", [[19, 20], [19, 20], [19, 20], [19, 20]]],
        ["g[5].partitioned_parpids[0]", [[5, 9], [6, 10], [7, 11], [8, 12]]],
    ],
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_compile.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print(f"wrote {out}")
