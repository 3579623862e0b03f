# Builds the reference's own host-side status decoder (decode_debuginfo_bicgstab
# and decode_debuginfo_bicgstab_query) from its sources, where they lie, into
# oracle/_ref/.  Used by tests only, to decode the status and limits lines this
# solver encodes.  Nothing of the reference is copied.
#   make -f oracle/ref.mk
REF ?= /root/reference/linearalgebra/ilu0bicgstab/xilinx/src/sda_app/common
OUT := oracle/_ref

$(OUT)/libbda_status_ref.so: $(REF)/bicgstab_utils.cpp $(REF)/bda_utils.cpp
	mkdir -p $(OUT)
	g++ -O2 -fPIC -shared -I$(REF) $(REF)/bicgstab_utils.cpp $(REF)/bda_utils.cpp -o $@
