# Host planner builds under sanitizers, on the CPU (g++ only; no GPU code):
#   _san/test_runs_asan        tests/cpp/test_runs.cpp, ASan + UBSan
#   _san/test_runs_tsan        the same with TSan (the learned planner's
#                              parallel candidate planning)
#   _san/test_batch_plan_asan  tests/cpp/test_batch_plan.cpp, ASan + UBSan
#   _san/test_batch_plan_tsan  the same with TSan (16-thread planning)
# The planner sources are compiled in directly; tests/cpp/san_glue.cpp stands
# in for the GPU model inference (no model, or a fixed stand-in function
# when KET_SAN_STUB_MODEL is set).
#   make -C ket_amd/cpp -f san.mk
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
ROOT := $(abspath $(HERE)../..)
CSRC := $(ROOT)/ket_amd/csrc
OUT := $(ROOT)/ket_amd/_san
CXX ?= g++
BASE := -std=c++20 -O1 -g -fno-omit-frame-pointer -ffp-contract=off -Wall \
  -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I$(ROOT)/include -I$(CSRC)
ASAN := -fsanitize=address,undefined -fno-sanitize-recover=all
TSAN := -fsanitize=thread
PLANNER := $(CSRC)/plan.cpp $(CSRC)/tuning.cpp $(ROOT)/tests/cpp/san_glue.cpp
DEPS := $(PLANNER) $(wildcard $(CSRC)/*.hpp) $(ROOT)/include/ket_amd.h

.PHONY: all clean
all: $(OUT)/test_runs_asan $(OUT)/test_runs_tsan $(OUT)/test_batch_plan_asan \
  $(OUT)/test_batch_plan_tsan

$(OUT)/test_runs_asan: $(ROOT)/tests/cpp/test_runs.cpp $(DEPS)
	@mkdir -p $(OUT)
	$(CXX) $(BASE) $(ASAN) -pthread -o $@ $< $(PLANNER)

$(OUT)/test_runs_tsan: $(ROOT)/tests/cpp/test_runs.cpp $(DEPS)
	@mkdir -p $(OUT)
	$(CXX) $(BASE) $(TSAN) -pthread -o $@ $< $(PLANNER)

$(OUT)/test_batch_plan_asan: $(ROOT)/tests/cpp/test_batch_plan.cpp $(CSRC)/batch_plan.cpp $(DEPS)
	@mkdir -p $(OUT)
	$(CXX) $(BASE) $(ASAN) -pthread -o $@ $< $(CSRC)/batch_plan.cpp $(PLANNER)

$(OUT)/test_batch_plan_tsan: $(ROOT)/tests/cpp/test_batch_plan.cpp $(CSRC)/batch_plan.cpp $(DEPS)
	@mkdir -p $(OUT)
	$(CXX) $(BASE) $(TSAN) -pthread -o $@ $< $(CSRC)/batch_plan.cpp $(PLANNER)

clean:
	rm -rf $(OUT)
