# Sanitizer builds of the CPU oracle (test infrastructure, CPU only; nothing
# here is built for or run on a GPU).  Outputs go to oracle/_san/:
#   oracle_asan  gcc, AddressSanitizer + UndefinedBehaviorSanitizer (libgomp)
#   oracle_tsan  clang, ThreadSanitizer with LLVM's libomp; run it with
#                OMP_TOOL_LIBRARIES=$(LLVM)/lib/libarcher.so so that TSan sees
#                OpenMP's barriers and critical sections (Archer, LLVM's OMPT tool)
# Usage: make -f oracle/sanitize.mk [asan|tsan]
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_san
LLVM ?= /opt/rocm/llvm
SRCS := $(HERE)vox_oracle.c $(HERE)vox_writers.c $(HERE)sanitize_main.c
COMMON := -O1 -g -std=gnu11 -fopenmp -ffp-contract=off -fno-fast-math -fno-omit-frame-pointer -Wall

all: asan tsan
asan: $(OUT)/oracle_asan
tsan: $(OUT)/oracle_tsan

$(OUT)/oracle_asan: $(SRCS)
	mkdir -p $(OUT)
	gcc $(COMMON) -fsanitize=address,undefined -fno-sanitize-recover=undefined -o $@ $(SRCS) -lm

$(OUT)/oracle_tsan: $(SRCS)
	mkdir -p $(OUT)
	$(LLVM)/bin/clang $(COMMON) -fsanitize=thread -o $@ $(SRCS) -lm -L$(LLVM)/lib -Wl,-rpath,$(LLVM)/lib

clean:
	rm -rf $(OUT)

.PHONY: all asan tsan clean
