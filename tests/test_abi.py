"""The C ABI: libmcsim.so loads and exports every entry point include/mcsim.h declares,
and the ctypes mirror matches the header's struct layouts (no GPU call is made)."""
import ctypes as C
import os
import re
import subprocess

from mcsim import _abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcsim.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mc_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    L = lib.load()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(lib.EXPORTS) == names


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "mcsim.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(mc_tool), sizeof(mc_mode), sizeof(mc_material),
         sizeof(mc_cell), sizeof(mc_trial), sizeof(mc_bisect_opts), sizeof(mc_limit), sizeof(mc_stamp));
  printf("%zu %zu %zu\n", offsetof(mc_cell, S), offsetof(mc_cell, revs), offsetof(mc_limit, steps));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(mc_stamp_opt), sizeof(mc_engage_query), sizeof(mc_engage_result),
         sizeof(mc_train_opts), sizeof(mc_lobe_args), sizeof(mc_search_tuning));
  printf("%zu %zu\n", offsetof(mc_stamp_opt, lut), offsetof(mc_engage_result, imm));
  printf("%zu %zu\n", sizeof(mc_engage_spec), offsetof(mc_engage_spec, Hb));
  printf("%zu %zu %zu %zu %zu\n", sizeof(mc_engine_info), sizeof(mc_wall), sizeof(mc_engine_state),
         sizeof(mc_engine_stats), sizeof(mc_chip_event));
  printf("%zu %zu %zu\n", offsetof(mc_engine_state, stepCount), offsetof(mc_engine_state, hist_i),
         offsetof(mc_engine_info, stepsPerTooth));
  printf("%zu %zu %zu %zu %zu\n", sizeof(mc_substep), sizeof(mc_substep_out), offsetof(mc_substep_out, arc),
         sizeof(mc_sim_program), offsetof(mc_sim_program, cornerR));
  printf("%zu %zu %zu %zu\n", sizeof(mc_frame_out), offsetof(mc_frame_out, end), offsetof(mc_substep, guard),
         offsetof(mc_sim_program, frames));
  printf("%zu %zu %zu %zu %zu\n", sizeof(mc_move), offsetof(mc_move, axes), offsetof(mc_move, sweep),
         sizeof(mc_path_point), sizeof(mc_schedule_opts));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(mc_wall_out), offsetof(mc_wall_out, markVal), offsetof(mc_wall_out, found),
         offsetof(mc_sim_program, flex), offsetof(mc_sim_program, kTool), offsetof(mc_sim_program, walls));
  return 0;
}'''
    import tempfile
    d = tempfile.mkdtemp()
    cpath, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
    open(cpath, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), cpath, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    want = [C.sizeof(t) for t in (_abi.Tool, _abi.Mode, _abi.Material, _abi.Cell, _abi.Trial,
                                  _abi.BisectOpts, _abi.Limit, _abi.Stamp)]
    want += [_abi.Cell.S.offset, _abi.Cell.revs.offset, _abi.Limit.steps.offset]
    want += [C.sizeof(t) for t in (_abi.StampOpt, _abi.EngageQuery, _abi.EngageResult, _abi.TrainOpts,
                                   _abi.LobeArgs, _abi.SearchTuning)]
    want += [_abi.StampOpt.lut.offset, _abi.EngageResult.imm.offset]
    want += [C.sizeof(_abi.EngageSpec), _abi.EngageSpec.Hb.offset]
    want += [C.sizeof(t) for t in (_abi.EngineInfo, _abi.Wall, _abi.EngineState, _abi.EngineStats, _abi.ChipEvent)]
    want += [_abi.EngineState.stepCount.offset, _abi.EngineState.hist_i.offset, _abi.EngineInfo.stepsPerTooth.offset]
    want += [C.sizeof(_abi.Substep), C.sizeof(_abi.SubstepOut), _abi.SubstepOut.arc.offset, C.sizeof(_abi.SimProgram),
             _abi.SimProgram.cornerR.offset]
    want += [C.sizeof(_abi.FrameOut), _abi.FrameOut.end.offset, _abi.Substep.guard.offset, _abi.SimProgram.frames.offset]
    want += [C.sizeof(_abi.Move), _abi.Move.axes.offset, _abi.Move.sweep.offset, C.sizeof(_abi.PathPoint),
             C.sizeof(_abi.ScheduleOpts)]
    want += [C.sizeof(_abi.WallOut), _abi.WallOut.markVal.offset, _abi.WallOut.found.offset,
             _abi.SimProgram.flex.offset, _abi.SimProgram.kTool.offset, _abi.SimProgram.walls.offset]
    assert [int(x) for x in out] == want
