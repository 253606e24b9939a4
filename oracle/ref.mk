# Builds the reference's own DCD reader (vmdplugins/dcdplugin.c, compiled from
# the reference sources where they lie) with the test harness dcd_ref.c into
# _ref/libdcdref.so.  Test infrastructure only; nothing is copied from the
# reference.  Plain -O2 for the host's baseline ISA: no FMA contraction, so
# the harness's double arithmetic rounds like erdst/vmdfio.c's.
REF ?= /root/reference
CC ?= gcc
OUT := _ref

all: $(OUT)/libdcdref.so

$(OUT)/libdcdref.so: dcd_ref.c $(REF)/vmdplugins/dcdplugin.c
	mkdir -p $(OUT)
	$(CC) -O2 -fPIC -shared -w -I$(REF)/vmdplugins -o $@ $(REF)/vmdplugins/dcdplugin.c dcd_ref.c -lm

clean:
	rm -rf $(OUT)
