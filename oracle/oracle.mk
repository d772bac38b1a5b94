# Test infrastructure only: builds the CPU restatement (the oracle) as a
# shared library for the ctypes wrapper in oracle/oracle.py.
#
#   make -f oracle/oracle.mk     -> oracle/liboracle.so
#
# -ffp-contract=off and the default x86-64 -march keep every a*b+c a separate
# multiply and add, as in the reference CPU build.

HERE := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
CC   := gcc
CFLAGS := -std=gnu99 -O2 -fPIC -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function -fopenmp

all: $(HERE)/liboracle.so

$(HERE)/liboracle.so: $(HERE)/mmca_oracle.c $(HERE)/../include/mmca.h
	$(CC) $(CFLAGS) -shared $(HERE)/mmca_oracle.c -o $@ -lm -lgomp

clean:
	rm -f $(HERE)/liboracle.so

.PHONY: all clean
