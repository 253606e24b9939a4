/*
 * Test infrastructure, not product code: reads DCD files through the
 * reference's own VMD DCD plugin (vmdplugins/dcdplugin.c, compiled from the
 * reference sources into oracle/_ref/ by oracle/ref.mk) and forms each frame's
 * cell as erdst/vmdfio.c:355-378 does, so the tests can compare the native
 * trajectory reader with what erdst itself reads.  Only tests/ load it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "molfile_plugin.h"

static molfile_plugin_t *dcd_plugin;

static int take_dcd(void *unused, vmdplugin_t *p)
{
    (void)unused;
    if (!strcmp(p->type, MOLFILE_PLUGIN_TYPE) && !strcmp(p->name, "dcd")) dcd_plugin = (molfile_plugin_t *)p;
    return VMDPLUGIN_SUCCESS;
}

/* Reads up to maxframes frames of path into xyz [frame][natoms][3] and cell
 * [frame][9] (column-major a, b, c).  Returns the number of frames read, or
 * -1 when the plugin cannot open the file; *natoms_out gets the atom count. */
int ref_dcd_read(const char *path, int maxframes, int natoms_cap, float *xyz, float *cell, int *natoms_out)
{
    if (!dcd_plugin) {
        vmdplugin_init();
        vmdplugin_register(NULL, take_dcd);
        if (!dcd_plugin) return -2;
    }
    int natoms = 0;
    void *h = dcd_plugin->open_file_read(path, "dcd", &natoms);
    if (!h) return -1;
    *natoms_out = natoms;
    if (natoms > natoms_cap) {
        dcd_plugin->close_file_read(h);
        return -3;
    }
    int n = 0;
    while (n < maxframes) {
        molfile_timestep_t ts;
        memset(&ts, 0, sizeof ts);
        ts.coords = xyz + (size_t)n * 3 * natoms;
        if (dcd_plugin->read_next_timestep(h, natoms, &ts) != MOLFILE_SUCCESS) break;
        /* erdst/vmdfio.c:365-378 */
        double x = cos(ts.gamma * M_PI / 180.0);
        double y = sin(ts.gamma * M_PI / 180.0);
        double u = cos(ts.beta * M_PI / 180.0);
        double v = (cos(ts.alpha * M_PI / 180.0) - x * u) / y;
        double w = sqrt(1 - u * u - v * v);
        float *b = cell + (size_t)9 * n;
        b[0] = ts.A;
        b[3] = ts.B * x;
        b[4] = ts.B * y;
        b[6] = ts.C * u;
        b[7] = ts.C * v;
        b[8] = ts.C * w;
        b[1] = 0.0f; b[2] = 0.0f; b[5] = 0.0f;
        n++;
    }
    dcd_plugin->close_file_read(h);
    return n;
}
